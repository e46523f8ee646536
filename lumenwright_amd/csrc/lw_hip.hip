// HIP backend of the HQ path tracer: the gfx950 kernels and the C ABI of include/lw/abi.h.
//
// Frame structure (one render_hq call), mirroring CudaBackend::render_hq
// (crates/lumenwright/src/render/cuda.rs:77-184) with the per-pixel sum in f64:
//   for each launch of `batch` samples, in sample order:
//     k_render:     one thread per (pixel, sample); a wave is an 8 x 8 pixel tile of one
//                   sample index, so primary rays of a wave are coherent. Each thread
//                   writes the sample's f32 radiance to a staging buffer.
//     k_accumulate: one thread per pixel adds the batch's finite samples, in sample order,
//                   to the pixel's f64 sums.
//   k_resolve:      mean = sum / spp, rounded to f32 (render/cpu/mod.rs:75-79).
// Because every pixel's f64 additions happen in the same order as the CPU backend's
// (sample 0, 1, 2, ...), the image equals the CPU backend's wherever the samples do.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <thread>
#include <string>
#include <vector>

#include "lw/abi.h"
#include "lw_device.hpp"
#include "lw_host.hpp"
#include "lw_path.hpp"
#include "lw_diff.hpp"
#include "lw_build.hpp"
#include "lw_sky.hpp"
#include "lw_quick_sky.hpp"
#include "lw_libm.hpp"
#include "lw_quick.hpp"

using namespace lw;

// The scratch memory of lw_hip_render_hq_scene_noalloc: a frame's f64 accumulator followed by
// its staged samples (and, once they are summed, its f32 means). Each device that loads the
// library holds its own copy until the process ends. In host code `scratch` names a host-side
// stand-in, never the device memory: kernels and copies take the address that
// hipGetSymbolAddress gives for the current device (g_scratch_ptr, set when a backend opens).
constexpr size_t kScratchBytes = LW_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxScratchDevices] = {};
}  // namespace

// ------------------------------------------------------------------------------------
// kernels

namespace {

constexpr int kBlock = 256;  // 4 waves of 64: a 16 x 16 pixel tile
// The persistent tile kernel's work item: a tile of one view's pixels with all their samples,
// rendered by one wave as 2^sh pixels x (64 >> sh) sample slots. The tile shape sh, one for a
// whole launch: 0 is one pixel (1 x 1), 1 two pixels side by side (2 x 1), 2 a 2 x 2 square.
constexpr __host__ __device__ uint32_t tile_w(uint32_t sh) { return sh ? 2u : 1u; }
constexpr __host__ __device__ uint32_t tile_h(uint32_t sh) { return sh == 2u ? 2u : 1u; }
__host__ __device__ __forceinline__ uint64_t tile_count(uint32_t w, uint32_t h, uint32_t sh) {
  return (uint64_t)((w + tile_w(sh) - 1u) / tile_w(sh)) * ((h + tile_h(sh) - 1u) / tile_h(sh));
}

__device__ __forceinline__ void load_lut(float *lut_s, const float *lut) {
  for (int i = threadIdx.x; i < 256; i += blockDim.x) lut_s[i] = lut[i];
  __syncthreads();
}

// Workgroups go to the 8 XCDs in turn (workgroup b to XCD b mod 8), each XCD with its own
// L2. xcd_tile renumbers them so that each XCD takes one contiguous run of the n tiles
// (neighbouring tiles trace through the same bricks): a bijection on [0, n).
__device__ __forceinline__ uint32_t xcd_tile(uint32_t b, uint32_t n) {
  const uint32_t q = n >> 3, r = n & 7u, xcd = b & 7u, idx = b >> 3;
  return (xcd < r ? xcd * (q + 1u) : r * (q + 1u) + (xcd - r) * q) + idx;
}

template <int W>
__global__ void __launch_bounds__(kBlock, W) k_render(const lw_params p, const DevScene s,
                                                   uint32_t sample_start, uint32_t xcd_remap,
                                                   float *__restrict__ rad) {
  __shared__ float lut_s[256];
  load_lut(lut_s, s.gamma_lut);
  const uint32_t w = p.frame[0], h = p.frame[1];
  const uint32_t tiles_x = (w + 15u) / 16u;
  const uint32_t tile = xcd_remap ? xcd_tile(blockIdx.x, gridDim.x) : blockIdx.x;
  const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
  const uint32_t px = (tile % tiles_x) * 16u + (wave & 1u) * 8u + (lane & 7u);
  const uint32_t py = (tile / tiles_x) * 16u + (wave >> 1) * 8u + (lane >> 3);
  if (px >= w || py >= h) return;
  const uint32_t sl = blockIdx.y;
  const V3 c = render_sample(s, p, lut_s, px, py, sample_start + sl);
  float *o = rad + ((size_t)sl * h * w + (size_t)py * w + px) * 3;
  o[0] = c.x;
  o[1] = c.y;
  o[2] = c.z;
}

// A launch lane's pixel and sample within the launch, in layout L. L 0: a workgroup is a
// 16 x 16 pixel tile (2 x 2 waves of 8 x 8) at sample blockIdx.y. L 1: a 4 x 4 pixel tile
// (2 x 2 waves of 2 x 2 pixels x 16 samples) at samples 16 blockIdx.y .. + 15, for launches of
// a multiple of 16 samples per pixel. L 2: a 2 x 2 pixel tile (a wave per pixel, its 64
// lanes 64 samples) at samples 64 blockIdx.y .. + 63, for launches of a multiple of 64; the
// wave's pixel is then the same for all its lanes (scalar registers). The layouts differ only
// in which lane renders a (pixel, sample), never in its value.
struct LaunchLane {
  uint32_t px, py, sl;
};
template <int L>
__device__ __forceinline__ LaunchLane launch_lane(uint32_t w) {
  const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63u;
  if constexpr (L == 2) {
    const uint32_t tx = (w + 1u) / 2u;
    return LaunchLane{(blockIdx.x % tx) * 2u + (wave & 1u), (blockIdx.x / tx) * 2u + (wave >> 1),
                      blockIdx.y * 64u + lane};
  } else if constexpr (L == 1) {
    const uint32_t tx = (w + 3u) / 4u, pix = lane & 3u;
    return LaunchLane{(blockIdx.x % tx) * 4u + (wave & 1u) * 2u + (pix & 1u),
                      (blockIdx.x / tx) * 4u + (wave >> 1) * 2u + (pix >> 1), blockIdx.y * 16u + (lane >> 2)};
  } else {
    const uint32_t tx = (w + 15u) / 16u;
    return LaunchLane{(blockIdx.x % tx) * 16u + (wave & 1u) * 8u + (lane & 7u),
                      (blockIdx.x / tx) * 16u + (wave >> 1) * 8u + (lane >> 3), blockIdx.y};
  }
}
// The layout for a launch of nb samples per pixel, and its grid.
__host__ __forceinline__ uint32_t launch_layout(uint32_t nb) {
  return nb % 64u == 0 ? 2u : (nb % 16u == 0 ? 1u : 0u);
}
__host__ __forceinline__ dim3 launch_grid(uint32_t lay, uint32_t w, uint32_t h, uint32_t nb) {
  if (lay == 2u) return dim3(((w + 1) / 2) * ((h + 1) / 2), nb / 64u);
  if (lay == 1u) return dim3(((w + 3) / 4) * ((h + 3) / 4), nb / 16u);
  return dim3(((w + 15) / 16) * ((h + 15) / 16), nb);
}

// k_render with 64-lane waves of 2 x 2 pixels x 16 samples, for launches of a multiple of 16
// samples per pixel: a workgroup is a 4 x 4 pixel tile (2 x 2 waves) at 16 consecutive samples
// of the launch (blockIdx.y picks which 16), so a wave's paths are 16 jittered samples of each
// of 4 neighbouring pixels, whose first walks and bounces stay close together (as in
// k_render_tiles). The staging is indexed as k_render's, so the launch's accumulation is
// unchanged.
template <int W, int L>
__global__ void __launch_bounds__(kBlock, W) k_render_slots(const lw_params p, const DevScene s,
                                                         uint32_t sample_start, float *__restrict__ rad) {
  __shared__ float lut_s[256];
  load_lut(lut_s, s.gamma_lut);
  const uint32_t w = p.frame[0], h = p.frame[1];
  const LaunchLane Ln = launch_lane<L>(w);
  const uint32_t px = Ln.px, py = Ln.py, sl = Ln.sl;
  if (px >= w || py >= h) return;
  const V3 c = render_sample(s, p, lut_s, px, py, sample_start + sl);
  float *o = rad + ((size_t)sl * h * w + (size_t)py * w + px) * 3;
  o[0] = c.x;
  o[1] = c.y;
  o[2] = c.z;
}

// The trace calls of samples [sample_start, sample_start + gridDim.y) of every pixel, as
// k_render renders them (the same lanes, the same paths): summed per workgroup in LDS, one
// 64-bit atomic a workgroup.
__global__ void __launch_bounds__(kBlock) k_count_traces(const lw_params p, const DevScene s,
                                                         uint32_t sample_start,
                                                         unsigned long long *__restrict__ total) {
  __shared__ float lut_s[256];
  __shared__ uint32_t sums[kBlock / 64];
  load_lut(lut_s, s.gamma_lut);
  const uint32_t w = p.frame[0], h = p.frame[1];
  const uint32_t tiles_x = (w + 15u) / 16u;
  const uint32_t tile = blockIdx.x;
  const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
  const uint32_t px = (tile % tiles_x) * 16u + (wave & 1u) * 8u + (lane & 7u);
  const uint32_t py = (tile / tiles_x) * 16u + (wave >> 1) * 8u + (lane >> 3);
  uint32_t n = 0;
  if (px < w && py < h) {
    SampleState S;
    sample_begin(S, p, px, py, sample_start + blockIdx.y);
    while (sample_event(S, s, p, lut_s, &n)) {
    }
  }
  for (uint32_t off = 32; off > 0; off >>= 1) n += __shfl_xor(n, off);
  if (lane == 0) sums[wave] = n;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    for (uint32_t k = 0; k < kBlock / 64; k++) t += sums[k];
    atomicAdd(total, t);
  }
}

__global__ void __launch_bounds__(kBlock) k_accumulate(uint32_t npix, uint32_t samples,
                                                       const float *__restrict__ rad,
                                                       double *__restrict__ acc) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npix) return;
  double a0 = acc[3 * (size_t)i], a1 = acc[3 * (size_t)i + 1], a2 = acc[3 * (size_t)i + 2];
  for (uint32_t sl = 0; sl < samples; sl++) {
    const float *c = rad + ((size_t)sl * npix + i) * 3;
    const float r = c[0], g = c[1], b = c[2];
    if (isfinite(r) && isfinite(g) && isfinite(b)) {
      a0 += (double)r;
      a1 += (double)g;
      a2 += (double)b;
    }
  }
  acc[3 * (size_t)i] = a0;
  acc[3 * (size_t)i + 1] = a1;
  acc[3 * (size_t)i + 2] = a2;
}

__global__ void __launch_bounds__(kBlock) k_zero(double *__restrict__ acc, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) acc[i] = 0.0;
}

__global__ void __launch_bounds__(kBlock) k_resolve(uint32_t npix, uint32_t spp,
                                                    const double *__restrict__ acc,
                                                    float *__restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 3u * npix) return;
  out[i] = (float)(acc[i] / (double)spp);
}

__global__ void __launch_bounds__(kBlock) k_trace(const lw_params p, const DevScene s, uint32_t n,
                                                  const float *__restrict__ o,
                                                  const float *__restrict__ d,
                                                  const float *__restrict__ tmin,
                                                  const float *__restrict__ tmax,
                                                  lw_hit *__restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  // The hit fields come from the walk the renderers run (no cell report, so it stops once
  // every brick left toward the ray's octant is empty); the final cell from a second, whole
  // walk.
  const V3 ro = V3{o[3 * i], o[3 * i + 1], o[3 * i + 2]}, rd = V3{d[3 * i], d[3 * i + 1], d[3 * i + 2]};
  Hit h, h_whole;
  int32_t cell[3];
  const bool hit = trace(s, p, ro, rd, tmin[i], tmax[i], h, nullptr);
  (void)trace(s, p, ro, rd, tmin[i], tmax[i], h_whole, cell);
  lw_hit r;
  r.t = hit ? h.t : INFINITY;
  r.quad = hit ? h.quad : UINT32_MAX;
  r.texel = hit ? h.texel : 0u;
  r.s = hit ? h.s : 0.0f;
  r.tt = hit ? h.tt : 0.0f;
  r.cell[0] = cell[0];
  r.cell[1] = cell[1];
  r.cell[2] = cell[2];
  r.hit = hit ? 1u : 0u;
  r.pad = 0u;
  out[i] = r;
}

__global__ void __launch_bounds__(kBlock) k_render_samples(const lw_params p, const DevScene s,
                                                           uint32_t n,
                                                           const uint32_t *__restrict__ px,
                                                           const uint32_t *__restrict__ py,
                                                           const uint32_t *__restrict__ sample,
                                                           float *__restrict__ out) {
  __shared__ float lut_s[256];
  load_lut(lut_s, s.gamma_lut);
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const V3 c = render_sample(s, p, lut_s, px[i], py[i], sample[i]);
  out[3 * i] = c.x;
  out[3 * i + 1] = c.y;
  out[3 * i + 2] = c.z;
}

// The samples of `samples` consecutive sample indices of a frame, as work items: item
// ((sl * tiles) + tile) * 256 + t is pixel t of 16 x 16 tile `tile` (four 8 x 8 wave
// tiles) at sample sample_start + sl.
struct FrameWork {
  uint32_t w, h, tiles_x, tiles, sample_start, total;
  float *__restrict__ rad;
  __device__ __forceinline__ bool decode(uint32_t item, uint32_t &px, uint32_t &py,
                                         uint32_t &sample) const {
    const uint32_t t = item & 255u, rest = item >> 8;
    const uint32_t tile = rest % tiles, sl = rest / tiles;
    const uint32_t wave = t >> 6, lane = t & 63u;
    px = (tile % tiles_x) * 16u + (wave & 1u) * 8u + (lane & 7u);
    py = (tile / tiles_x) * 16u + (wave >> 1) * 8u + (lane >> 3);
    sample = sample_start + sl;
    return px < w && py < h;
  }
  __device__ __forceinline__ uint32_t count() const { return total; }
  __device__ __forceinline__ bool start(const Tracer &T, PathState &P, uint32_t item) const {
    uint32_t px, py, sample;
    if (!decode(item, px, py, sample)) return false;
    T.start(P, T.p, px, py, sample);
    return true;
  }
  __device__ __forceinline__ void write(uint32_t item, V3 c) const {
    uint32_t px, py, sample;
    decode(item, px, py, sample);
    float *o = rad + ((size_t)(sample - sample_start) * h * w + (size_t)py * w + px) * 3;
    o[0] = c.x;
    o[1] = c.y;
    o[2] = c.z;
  }
};

// An explicit list of (px, py, sample) triples.
struct ListWork {
  const uint32_t *__restrict__ px;
  const uint32_t *__restrict__ py;
  const uint32_t *__restrict__ sample;
  float *__restrict__ out;
  uint32_t total;
  __device__ __forceinline__ bool decode(uint32_t item, uint32_t &x, uint32_t &y,
                                         uint32_t &s) const {
    x = px[item];
    y = py[item];
    s = sample[item];
    return true;
  }
  __device__ __forceinline__ uint32_t count() const { return total; }
  __device__ __forceinline__ bool start(const Tracer &T, PathState &P, uint32_t item) const {
    T.start(P, T.p, px[item], py[item], sample[item]);
    return true;
  }
  __device__ __forceinline__ void write(uint32_t item, V3 c) const {
    out[3 * (size_t)item] = c.x;
    out[3 * (size_t)item + 1] = c.y;
    out[3 * (size_t)item + 2] = c.z;
  }
};

// ---- batches: many views of one scene in one work pool ----
//
// A batch renders its views in passes; pass k renders samples [k m_v, (k + 1) m_v) of view
// v (m_v = its samples per pass, clipped to its spp), so every pixel's samples still reach
// its f64 sum in sample order. k_batch_plan, one workgroup, turns the pass index kept in
// device memory into each view's sample range and the prefix sums of the pass's work items,
// so a pass is the same four launches every time and replays as one captured hipGraph.
constexpr uint32_t kMaxViews = 256;

struct BatchDev {
  ViewCam cam[kMaxViews];
  uint32_t tiles_x[kMaxViews], tiles[kMaxViews], per_pass[kMaxViews];
  uint64_t rad_base[kMaxViews];  // the view's staging, in floats from the buffer start
  uint32_t pix_prefix[kMaxViews + 1];
  uint32_t nviews, pass, total, npix_total;
  // Written by k_batch_plan for the current pass.
  uint32_t n[kMaxViews], start[kMaxViews], item_prefix[kMaxViews + 1];
};

// The largest v with prefix[v] <= x (prefix non-decreasing, prefix[0] = 0 <= x).
__device__ __forceinline__ uint32_t find_view(const uint32_t *prefix, uint32_t nviews, uint32_t x) {
  uint32_t lo = 0, hi = nviews;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (prefix[mid] <= x)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

struct BatchWork {
  const BatchDev *__restrict__ bd;
  float *__restrict__ rad;
  __device__ __forceinline__ uint32_t count() const { return bd->total; }
  __device__ __forceinline__ bool decode(uint32_t item, uint32_t &v, uint32_t &px, uint32_t &py,
                                         uint32_t &sl) const {
    v = find_view(bd->item_prefix, bd->nviews, item);
    const uint32_t local = item - bd->item_prefix[v];
    const uint32_t t = local & 255u, rest = local >> 8;
    const uint32_t tiles = bd->tiles[v], tx = bd->tiles_x[v];
    const uint32_t tile = rest % tiles;
    sl = rest / tiles;
    const uint32_t wave = t >> 6, lane = t & 63u;
    px = (tile % tx) * 16u + (wave & 1u) * 8u + (lane & 7u);
    py = (tile / tx) * 16u + (wave >> 1) * 8u + (lane >> 3);
    return px < bd->cam[v].width && py < bd->cam[v].height;
  }
  __device__ __forceinline__ bool start(const Tracer &T, PathState &P, uint32_t item) const {
    uint32_t v, px, py, sl;
    if (!decode(item, v, px, py, sl)) return false;
    const ViewCam cam = bd->cam[v];
    T.start(P, cam, px, py, bd->start[v] + sl);
    return true;
  }
  __device__ __forceinline__ void write(uint32_t item, V3 c) const {
    uint32_t v, px, py, sl;
    decode(item, v, px, py, sl);
    const uint32_t w = bd->cam[v].width, h = bd->cam[v].height;
    float *o = rad + bd->rad_base[v] + ((size_t)sl * h * w + (size_t)py * w + px) * 3;
    o[0] = c.x;
    o[1] = c.y;
    o[2] = c.z;
  }
};

// One workgroup of kMaxViews lanes: the pass's sample ranges and item prefix sums, then the
// pass index moves on.
__global__ void __launch_bounds__(kMaxViews) k_batch_plan(BatchDev *__restrict__ bd) {
  __shared__ uint32_t scan[kMaxViews];
  const uint32_t v = threadIdx.x;
  const uint32_t nv = bd->nviews, pass = bd->pass;
  uint32_t items = 0;
  if (v < nv) {
    const uint32_t spp = bd->cam[v].spp, m = bd->per_pass[v];
    const uint64_t s0 = (uint64_t)pass * m;
    const uint32_t n = s0 < spp ? (uint32_t)std::min<uint64_t>(m, spp - s0) : 0u;
    bd->n[v] = n;
    bd->start[v] = (uint32_t)std::min<uint64_t>(s0, spp);
    items = bd->tiles[v] * 256u * n;
  }
  scan[v] = items;
  __syncthreads();
  for (uint32_t off = 1; off < kMaxViews; off <<= 1) {
    const uint32_t add = v >= off ? scan[v - off] : 0u;
    __syncthreads();
    scan[v] += add;
    __syncthreads();
  }
  if (v < nv) bd->item_prefix[v] = scan[v] - items;
  if (v + 1 == nv) {
    bd->item_prefix[nv] = scan[v];
    bd->total = scan[v];
  }
  __syncthreads();
  if (v == 0) bd->pass = pass + 1u;
}

__global__ void __launch_bounds__(kBlock) k_accumulate_batch(const BatchDev *__restrict__ bd,
                                                             const float *__restrict__ rad,
                                                             double *__restrict__ acc) {
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= bd->npix_total) return;
  const uint32_t v = find_view(bd->pix_prefix, bd->nviews, g);
  const uint32_t i = g - bd->pix_prefix[v];
  const size_t npix = (size_t)bd->cam[v].width * bd->cam[v].height;
  const uint32_t n = bd->n[v];
  double a0 = acc[3 * (size_t)g], a1 = acc[3 * (size_t)g + 1], a2 = acc[3 * (size_t)g + 2];
  for (uint32_t sl = 0; sl < n; sl++) {
    const float *c = rad + bd->rad_base[v] + ((size_t)sl * npix + i) * 3;
    const float r = c[0], gg = c[1], bb = c[2];
    if (isfinite(r) && isfinite(gg) && isfinite(bb)) {
      a0 += (double)r;
      a1 += (double)gg;
      a2 += (double)bb;
    }
  }
  acc[3 * (size_t)g] = a0;
  acc[3 * (size_t)g + 1] = a1;
  acc[3 * (size_t)g + 2] = a2;
}

__global__ void __launch_bounds__(kBlock) k_resolve_batch(const BatchDev *__restrict__ bd,
                                                          const double *__restrict__ acc,
                                                          float *__restrict__ out) {
  const uint32_t g3 = blockIdx.x * blockDim.x + threadIdx.x;
  if (g3 >= 3u * bd->npix_total) return;
  const uint32_t v = find_view(bd->pix_prefix, bd->nviews, g3 / 3u);
  out[g3] = (float)(acc[g3] / (double)bd->cam[v].spp);
}

// The direct form of a batch pass: one workgroup per 256 work items (a 16 x 16 tile of one
// view at one sample), one lane per sample. Every view's items are a multiple of 256, so a
// workgroup lies in one view; workgroups past the pass's items return at once (the grid is
// sized for the first pass, the largest).
template <int W>
__global__ void __launch_bounds__(kBlock, W) k_render_batch(const lw_params p, const DevScene s,
                                                         const BatchDev *__restrict__ bd,
                                                         float *__restrict__ rad) {
  const uint32_t base = blockIdx.x * (uint32_t)kBlock;
  if (base >= bd->total) return;
  __shared__ float lut_s[256];
  load_lut(lut_s, s.gamma_lut);
  const uint32_t v = find_view(bd->item_prefix, bd->nviews, base);
  const uint32_t rest = (base - bd->item_prefix[v]) >> 8;
  const uint32_t tiles = bd->tiles[v], tx = bd->tiles_x[v];
  const uint32_t tile = rest % tiles, sl = rest / tiles;
  const ViewCam cam = bd->cam[v];
  const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
  const uint32_t px = (tile % tx) * 16u + (wave & 1u) * 8u + (lane & 7u);
  const uint32_t py = (tile / tx) * 16u + (wave >> 1) * 8u + (lane >> 3);
  if (px >= cam.width || py >= cam.height) return;
  const V3 c = render_sample(s, p, cam, lut_s, px, py, bd->start[v] + sl);
  float *o = rad + bd->rad_base[v] +
             ((size_t)sl * cam.height * cam.width + (size_t)py * cam.width + px) * 3;
  o[0] = c.x;
  o[1] = c.y;
  o[2] = c.z;
}

// The direct form in one persistent launch, for a frame or a batch of views. A work item is one
// tile of one view with all of its samples (view << 24 | tile, tiles numbered row by row;
// without `items`, item k is tile k of view 0), in the launch's tile shape SH (tile_w, tile_h).
// Whole waves take items from a device counter. A wave's lanes are the tile's 2^SH pixels x
// (64 >> SH) sample slots: in round r, slot j renders sample (64 >> SH) r + j of its pixel, so
// a round's paths are jittered samples of one pixel (or of 2 or 4 neighbouring pixels), whose
// first walks and first bounces stay close together. The lanes leave their results in the
// wave's row of LDS, and the pixel's slot-0 lane adds the finite ones to the pixel's f64 sums
// in sample order (the CPU backend's order, so the image is render_hq's bit for bit,
// render/cpu/mod.rs:66-79), kept in LDS; it writes the pixel's mean at the end. No staging and
// no accumulation pass. A wave leaves once the counter has passed the last item, so every wave
// ends. With `progress_host`, the index of each item taken is stored to host memory for the
// caller's progress reports.
// One item of k_render_tiles: the tile `tile` of view v in shape SH (tile_w, tile_h), all of
// its samples, its means written to `out`.
template <int SH>
__device__ __forceinline__ void tiles_item(const lw_params &p, const DevScene &s,
                                           const BatchDev *__restrict__ bd, const float *lut_s,
                                           double *sums, float4 *row, uint32_t v, uint32_t tile,
                                           float *__restrict__ out) {
  constexpr uint32_t tw = tile_w(SH), th = tile_h(SH);
  constexpr uint32_t np = 1u << SH, ns = 64u >> SH;  // pixels, and sample slots per pixel
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t pix = lane & (np - 1u), slot = lane >> SH;
  double *acc = sums + (size_t)((threadIdx.x >> 6) * 4u + pix) * 3;
  const ViewCam &cam = bd->cam[v];
  const uint32_t tiles_x = (cam.width + tw - 1u) / tw;
  const uint32_t px = (tile % tiles_x) * tw + pix % tw;
  const uint32_t py = (tile / tiles_x) * th + pix / tw;
  const bool inside = px < cam.width && py < cam.height;
  if (slot == 0u) {
    acc[0] = 0.0;
    acc[1] = 0.0;
    acc[2] = 0.0;
  }
  for (uint32_t base = 0; base < cam.spp; base += ns) {
    const uint32_t smp = base + slot;
    float4 r = make_float4(0.0f, 0.0f, 0.0f, 0.0f);  // w: 1 for a finite sample
    if (inside && smp < cam.spp) {
      const V3 c = render_sample(s, p, cam, lut_s, px, py, smp);
      if (isfinite(c.x) && isfinite(c.y) && isfinite(c.z)) r = make_float4(c.x, c.y, c.z, 1.0f);
    }
    row[lane] = r;
    __builtin_amdgcn_wave_barrier();
    if (slot == 0u && inside) {
      double a0 = acc[0], a1 = acc[1], a2 = acc[2];
      for (uint32_t j = 0; j < ns; j++) {
        const float4 q = row[pix + np * j];
        if (q.w != 0.0f) {
          a0 += (double)q.x;
          a1 += (double)q.y;
          a2 += (double)q.z;
        }
      }
      acc[0] = a0;
      acc[1] = a1;
      acc[2] = a2;
    }
    __builtin_amdgcn_wave_barrier();
  }
  if (slot == 0u && inside) {
    float *o = out + ((size_t)bd->pix_prefix[v] + (size_t)py * cam.width + px) * 3;
    o[0] = (float)(acc[0] / (double)cam.spp);
    o[1] = (float)(acc[1] / (double)cam.spp);
    o[2] = (float)(acc[2] / (double)cam.spp);
  }
}

// SH is the tile shape of every item, the same for all views of a launch. A fixed shape makes a
// one-pixel tile's pixel the same for the whole wave, which frees vector registers (168 B of
// spill per lane for one pixel against 196 for 2 x 2): 1147 against 1107 Msamples/s at 720p
// and 256 spp. The shape read per item (one copy of the code for every shape) ran as fast as
// 2 x 2 everywhere (1104), and one copy of the item's code per shape, picked per item, spilled
// more (232 B).
template <int W, int SH>
__global__ void __launch_bounds__(kBlock, W) k_render_tiles(const lw_params p, const DevScene s,
                                                  const BatchDev *__restrict__ bd,
                                                  const uint32_t *__restrict__ items, uint32_t n_items,
                                                  uint32_t *__restrict__ counter, float *__restrict__ out,
                                                  uint32_t *progress_host) {
  __shared__ float lut_s[256];
  __shared__ double sums[(kBlock / 64) * 4 * 3];  // up to 4 pixels per wave
  __shared__ float4 res[kBlock];
  load_lut(lut_s, s.gamma_lut);
  const uint32_t lane = threadIdx.x & 63u;
  float4 *row = res + (threadIdx.x & ~63u);  // this wave's results, one per lane
  for (;;) {
    uint32_t k = 0;
    if (lane == 0) {
      k = atomicAdd(counter, 1u);
      if (progress_host && k < n_items)
        __hip_atomic_store(progress_host, k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    }
    // The item and its view are the same for the whole wave: read into scalar registers, so
    // the view's camera is read by scalar loads where it is used rather than held per lane.
    k = __builtin_amdgcn_readfirstlane(__shfl(k, 0));
    if (k >= n_items) break;
    const uint32_t item = __builtin_amdgcn_readfirstlane(items ? items[k] : k);
    const uint32_t v = item >> 24, tile = item & 0xFFFFFFu;
    tiles_item<SH>(p, s, bd, lut_s, sums, row, v, tile, out);
  }
}

// ---- sky preparation (render/sky.rs, integrator.rs:133-148) ----

// build_sky_cache (integrator.rs:133-148), one lane per entry.
__global__ void __launch_bounds__(kBlock) k_sky_cache(const lw_params p, uint32_t n, float *__restrict__ out) {
  const uint32_t m = n + 1u;
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= m * m) return;
  lw::sky_cache_entry(p, n, k / m, k % m, out + (size_t)k * 3);
}

// blend_chunky_sky (sky.rs:836-851) over one w x h map, one lane per texel: the texel's
// direction (map_direction, sky.rs:856-860) from the per-column (cos, sin) of its azimuth and
// the per-row (cos, sin) of its height, and Chunky's sky there as the kernel reads it through
// the cache (sky_radiance).
__global__ void __launch_bounds__(kBlock) k_blend_sky(float *__restrict__ map, uint32_t w, uint32_t h,
                                                      const float2 *__restrict__ az,
                                                      const float2 *__restrict__ el,
                                                      const float *__restrict__ cache, const lw_params p,
                                                      float weight) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= w * h) return;
  const float2 a = az[k % w], e = el[k / w];
  const V3 d = v3(e.x * a.x, e.y, e.x * a.y);
  DevScene s{};
  s.sky = cache;
  const V3 c = sky_radiance(s, p, d);
  float *t = map + (size_t)k * 3;
  t[0] = weight * c.x + (1.0f - weight) * t[0];
  t[1] = weight * c.y + (1.0f - weight) * t[1];
  t[2] = weight * c.z + (1.0f - weight) * t[2];
}

// ---- post-processing (render/post.rs) ----

// A tone curve's inputs: the byte thresholds of the display encoding (host-made with libm's
// powf: the byte of v is the number of thresholds <= v), and per operator its Hable curve
// (a, b, c, d, e, f, w) and white scale.
struct ToneDev {
  float thr[255];
  float exposure, weight;
  uint32_t op_a, op_b;
  float ca[7], cb[7];
  double white_a, white_b;
};

// Rust std's f32::clamp: a NaN stays NaN (unlike the kernel crate's select-based clamp).
__device__ __forceinline__ float clamp_std(float x, float lo, float hi) {
  return x < lo ? lo : (x > hi ? hi : x);
}

// post.rs:23-26
__device__ __forceinline__ float aces_f(float x) {
  const float a = 2.51f, b = 0.03f, c = 2.43f, d = 0.59f, e = 0.14f;
  return clamp_std((x * (a * x + b)) / (x * (c * x + d) + e), 0.0f, 1.0f);
}

// post.rs:85-96: h(y) with the parameter products in f32 and the rest in f64.
__device__ __forceinline__ double hable_h(const float *k, double y) {
  const double a = (double)k[0], b = (double)k[1];
  const double cb = (double)(k[2] * k[1]), de = (double)(k[3] * k[4]), df = (double)(k[3] * k[5]),
               ef = (double)(k[4] / k[5]);
  return (y * (a * y + cb) + de) / (y * (a * y + b) + df) - ef;
}

// The operator's value before the display gamma (post.rs:155-165).
__device__ __forceinline__ float tone_linear(uint32_t op, const float *k, double white, float v,
                                             float exposure) {
  float x = v * exposure;
  x = isnan(x) ? 0.0f : (x > 0.0f ? x : 0.0f);  // f32::max(x, 0.0): the non-NaN operand
  if (op == LW_TONEMAP_GAMMA) return x;
  if (op == LW_TONEMAP_ACES) return aces_f(x);
  return (float)(hable_h(k, 2.0 * (double)x) * white);
}

// The display byte of v (post.rs:165): the number of thresholds <= v (NaN: none).
__device__ __forceinline__ uint32_t tone_byte(const float *thr, float v) {
  uint32_t c = 0;
#pragma unroll
  for (uint32_t step = 128; step > 0; step >>= 1)
    if (c + step <= 255u && thr[c + step - 1] <= v) c += step;
  return c;
}

__device__ __forceinline__ void load_thr(float *thr_s, const ToneDev &t) {
  for (int i = threadIdx.x; i < 255; i += blockDim.x) thr_s[i] = t.thr[i];
  __syncthreads();
}

// tonemap (post.rs:148-170), one lane per pixel.
__global__ void __launch_bounds__(kBlock) k_tonemap(const float *__restrict__ hdr, uint32_t n,
                                                    const ToneDev t, uint32_t *__restrict__ out) {
  __shared__ float thr_s[256];
  load_thr(thr_s, t);
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t px = 0xFF000000u;
#pragma unroll
  for (int c = 0; c < 3; c++)
    px |= tone_byte(thr_s, tone_linear(t.op_a, t.ca, t.white_a, hdr[3 * (size_t)i + c], t.exposure))
          << (8 * c);
  out[i] = px;
}

// The libm tables (lw_libm.hpp): powf(x, 1/2.2f), the display encoding (post.rs:141-144);
// powf(x, 2.2f), the texel decoding (sky.rs:872); log2f, the quick sampler's mip level
// (texture.rs:113).
constexpr float kPowExponent[2] = {1.0f / 2.2f, 2.2f};
enum { POW_ENCODE = 0, POW_DECODE = 1, LIBM_LOG2 = 2 };
using lw::pow_band;
using lw::pow_libm;

__device__ __forceinline__ float encode_f(float v, const uint2 *__restrict__ fix, uint32_t nfix) {
  return pow_libm(clamp_std(v, 0.0f, 1.0f), kPowExponent[POW_ENCODE], fix, nfix);
}

// The band inputs among the f32 bit patterns [u0, u0 + n) with their rounded f64 values of
// powf(x, y) or log2f(x), appended to out (one atomic per wave; entries past cap are counted,
// not written).
__global__ void __launch_bounds__(kBlock) k_pow_band(uint32_t u0, uint32_t n, int which, float y,
                                                     uint2 *__restrict__ out, uint32_t *__restrict__ count,
                                                     uint32_t cap) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t stride = gridDim.x * blockDim.x;
  // Every lane of a wave runs the same trip count, so the ballot sees the whole wave.
  const uint32_t base0 = (blockIdx.x * blockDim.x + threadIdx.x) & ~63u;
  for (uint32_t w0 = base0; w0 < n; w0 += stride) {
    const uint32_t u = u0 + w0 + lane;
    bool in_band = false;
    float c = 0.0f;
    if (w0 + lane < n)
      c = which == LIBM_LOG2 ? lw::log2_band(__uint_as_float(u), in_band) : pow_band(__uint_as_float(u), y, in_band);
    const uint64_t m = __ballot(in_band);
    if (m == 0) continue;
    uint32_t base = 0;
    if (lane == 0) base = atomicAdd(count, (uint32_t)__popcll(m));
    base = __shfl(base, 0);
    if (in_band) {
      const uint32_t slot = base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
      if (slot < cap) out[slot] = make_uint2(u, __float_as_uint(c));
    }
  }
}

// tonemap_mix (post.rs:176-215) for 0 < weight < 1.
__global__ void __launch_bounds__(kBlock) k_tonemap_mix(const float *__restrict__ hdr, uint32_t n,
                                                        const ToneDev t, const uint2 *__restrict__ fix,
                                                        uint32_t nfix, uint32_t *__restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t px = 0xFF000000u;
#pragma unroll
  for (int c = 0; c < 3; c++) {
    const float v = hdr[3 * (size_t)i + c];
    const float d = t.weight * encode_f(tone_linear(t.op_a, t.ca, t.white_a, v, t.exposure), fix, nfix) +
                    (1.0f - t.weight) * encode_f(tone_linear(t.op_b, t.cb, t.white_b, v, t.exposure), fix, nfix);
    const float b = d * 255.0f + 0.5f;
    px |= (b > 0.0f ? (uint32_t)b : 0u) << (8 * c);
  }
  out[i] = px;
}

// The display encoding of n values (post.rs:141-144), as the mix computes it.
__global__ void __launch_bounds__(kBlock) k_encode(const float *__restrict__ x, uint32_t n,
                                                   const uint2 *__restrict__ fix, uint32_t nfix,
                                                   float *__restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = encode_f(x[i], fix, nfix);
}

// Rust's f32 `as i32` and `as u32` of n values, as every kernel converts them.
__global__ void __launch_bounds__(kBlock) k_casts(const float *__restrict__ x, uint32_t n,
                                                  int32_t *__restrict__ out_i, uint32_t *__restrict__ out_u) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out_i[i] = lw::as_i32(x[i]);
  out_u[i] = lw::as_u32(x[i]);
}

// draw_map (sky.rs:864-881) for the game's maps, one lane per texel: the client's sky along
// the texel's direction, decoded to linear as c^2.2 with libm's powf (pow_libm).
__global__ void __launch_bounds__(kBlock) k_quick_sky_map(float *__restrict__ out, uint32_t w, uint32_t h,
                                                          const float2 *__restrict__ az,
                                                          const float2 *__restrict__ el,
                                                          const lw_quick_sky q, const lw::DiscSectors sec,
                                                          const uint32_t *__restrict__ atlas, uint32_t atlas_w,
                                                          uint32_t atlas_len, const uint32_t *__restrict__ stars,
                                                          const uint2 *__restrict__ fix, uint32_t nfix) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= w * h) return;
  const float2 a = az[k % w], e = el[k / w];
  const V3 d = v3(e.x * a.x, e.y, e.x * a.y);
  const V3 c = lw::qs_background(q, sec, d, atlas, atlas_w, atlas_len, stars);
  const float y = kPowExponent[POW_DECODE];
  float *t = out + (size_t)k * 3;
  t[0] = pow_libm(c.x, y, fix, nfix);
  t[1] = pow_libm(c.y, y, fix, nfix);
  t[2] = pow_libm(c.z, y, fix, nfix);
}

// The light's map scaled with its veil (sky.rs:950): (c + veil) * scale per texel.
__global__ void __launch_bounds__(kBlock) k_sky_veil(float *__restrict__ map, uint32_t n, float3 veil,
                                                     float scale) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  float *t = map + (size_t)k * 3;
  t[0] = (t[0] + veil.x) * scale;
  t[1] = (t[1] + veil.y) * scale;
  t[2] = (t[2] + veil.z) * scale;
}

// downsample (post.rs:218-247), one lane per output pixel.
__global__ void __launch_bounds__(kBlock) k_downsample(const uint32_t *__restrict__ px, uint32_t w,
                                                       uint32_t k, uint32_t ow, uint32_t oh,
                                                       uint32_t *__restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ow * oh) return;
  const uint32_t x = i % ow, y = i / ow;
  uint32_t acc[4] = {0u, 0u, 0u, 0u};
  for (uint32_t dy = 0; dy < k; dy++)
    for (uint32_t dx = 0; dx < k; dx++) {
      const uint32_t c = px[(size_t)(y * k + dy) * w + x * k + dx];
#pragma unroll
      for (int j = 0; j < 4; j++) acc[j] += (c >> (8 * j)) & 0xFFu;
    }
  const uint32_t nk = k * k;
  uint32_t v = 0;
#pragma unroll
  for (int j = 0; j < 4; j++) v += ((acc[j] + nk / 2) / nk) << (8 * j);
  out[i] = v;
}

// A plain streaming copy, 16 B per lane per step: the reference point that turns hardware
// byte counters into bytes for the path kernels.
__global__ void __launch_bounds__(kBlock) k_copy(const uint4 *__restrict__ src, uint4 *__restrict__ dst,
                                                 size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

__global__ void __launch_bounds__(kBlock) k_fill_u32(uint32_t *__restrict__ p, uint32_t v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

// DevScene::cell_ext, one workgroup per stored brick. For each ray octant k (sign sx, sy, sz
// per axis), the side of the largest model-free cube from a voxel toward k, clipped to the
// brick, follows n(v) = 0 for a voxel with a model, else min(8, 1 + min n(v + forward
// neighbour)) over the (up to 7) forward neighbours inside the brick (8 where there are none).
// In coordinates mirrored so that k points toward -1 on every axis, a voxel's forward
// neighbours lie on the planes x' + y' + z' below its own, so the planes are filled in order,
// one lane per (x, z) column, with a barrier between planes.
__global__ void __launch_bounds__(kBlock) k_cell_ext(const uint32_t *__restrict__ grid, uint64_t off_bricks,
                                                    uint32_t *__restrict__ ext) {
  __shared__ uint8_t n[4096];
  __shared__ uint32_t acc[4096];
  const uint32_t *words = grid + off_bricks + (uint64_t)blockIdx.x * LW_BRICK_WORDS;
  const int32_t cx = (int32_t)(threadIdx.x & 15u), cz = (int32_t)(threadIdx.x >> 4);
  uint32_t full = 0;  // bit y: voxel (cx, y, cz) has a model
  for (int32_t y = 0; y < 16; y++) {
    const uint32_t local = (uint32_t)((y * 16 + cz) * 16 + cx);
    if ((words[local >> 1] >> ((local & 1u) * 16u)) & 0xFFFFu) full |= 1u << y;
    acc[local] = 0u;
  }
  for (uint32_t k = 0; k < 8; k++) {
    const int32_t sx = (k & 1u) ? 1 : -1, sy = (k & 2u) ? 1 : -1, sz = (k & 4u) ? 1 : -1;
    const int32_t mx = sx > 0 ? 15 - cx : cx, mz = sz > 0 ? 15 - cz : cz;
    for (int32_t d = 0; d <= 45; d++) {
      const int32_t my = d - mx - mz;
      if (my >= 0 && my <= 15) {
        const int32_t y = sy > 0 ? 15 - my : my;
        const uint32_t local = (uint32_t)((y * 16 + cz) * 16 + cx);
        uint32_t m = 0;
        if (!((full >> y) & 1u)) {
          m = 8;
          for (int32_t c = 1; c < 8; c++) {
            const int32_t nx = cx + ((c & 1) ? sx : 0), ny = y + ((c & 2) ? sy : 0), nz = cz + ((c & 4) ? sz : 0);
            if (nx >= 0 && nx <= 15 && ny >= 0 && ny <= 15 && nz >= 0 && nz <= 15)
              m = min(m, (uint32_t)n[(ny * 16 + nz) * 16 + nx]);
          }
          m = min(m + 1u, 8u);
        }
        n[local] = (uint8_t)m;
        acc[local] |= m << (4u * k);
      }
      __syncthreads();
    }
  }
  uint32_t *out = ext + (uint64_t)blockIdx.x * 4096u;
  for (uint32_t i = threadIdx.x; i < 4096u; i += blockDim.x) out[i] = acc[i];
}

// DevScene::octant_boxes, in three kernels: k_octants_init marks each brick that is not
// EMPTY (a top-level or overlay entry) with all eight bits; k_octants_scan then ORs the bits
// along every line of bricks on one axis, toward +axis for the octants whose bit for that
// axis is set and toward -axis for the others. After the x, y and z passes, bit k of a brick
// is the OR over the box toward octant k. Bricks are numbered (y * dz + z) * dx + x.
__global__ void __launch_bounds__(kBlock) k_octants_init(const uint32_t *__restrict__ grid, uint32_t off_top,
                                                        uint32_t off_ov, uint64_t nb,
                                                        uint8_t *__restrict__ boxes) {
  for (uint64_t b = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; b < nb; b += (uint64_t)gridDim.x * blockDim.x)
    boxes[b] = (grid[off_top + b] != LW_EMPTY || grid[off_ov + b] != LW_EMPTY) ? 0xFFu : 0u;
}

// Lines along `axis` (0 x, 1 y, 2 z), one per lane, grid-stride.
__global__ void __launch_bounds__(kBlock) k_octants_scan(uint8_t *__restrict__ boxes, uint32_t dx, uint32_t dy,
                                                        uint32_t dz, uint32_t axis) {
  uint32_t plus = 0;
  for (int k = 0; k < 8; k++)
    if ((k >> axis) & 1) plus |= 1u << k;
  const uint32_t minus = 0xFFu & ~plus;
  const uint64_t lines = axis == 0 ? (uint64_t)dy * dz : (axis == 1 ? (uint64_t)dz * dx : (uint64_t)dy * dx);
  for (uint64_t line = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; line < lines;
       line += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t base, stride, len;
    if (axis == 0) {  // lines (y, z): base (y * dz + z) * dx
      base = line * dx, stride = 1, len = dx;
    } else if (axis == 1) {  // lines (z, x): base z * dx + x
      base = line, stride = (uint64_t)dx * dz, len = dy;
    } else {  // lines (y, x): base y * dz * dx + x
      base = (line / dx) * dz * dx + line % dx, stride = dx, len = dz;
    }
    uint32_t run = 0;
    for (uint64_t i = len; i-- > 0;) {  // toward +axis: the OR of this brick and those after it
      uint8_t &v = boxes[base + i * stride];
      run |= v & plus;
      v = (uint8_t)((v & minus) | run);
    }
    run = 0;
    for (uint64_t i = 0; i < len; i++) {  // toward -axis: the OR of this brick and those before it
      uint8_t &v = boxes[base + i * stride];
      run |= v & minus;
      v = (uint8_t)((v & plus) | run);
    }
  }
}

// Persistent path tracing: each lane runs one path at a time as a PathState; the wave
// traces one ray per lane per iteration, whichever kind of ray each lane needs. A lane
// whose path ends writes its radiance and takes the next work item: the wave's idle lanes
// are counted by a ballot, one atomic per wave reserves that many items, and each lane
// takes its rank among them. A wave ends once the items are exhausted and all its lanes
// are done; every path ends (MAX_EVENTS, MAX_STEPS), so every wave ends.
// Gives every idle lane of the wave a path: the wave's idle lanes are counted by a ballot,
// one atomic reserves that many work items, and each lane takes its rank among them
// (mbcnt). A lane whose item is outside the frame asks again; a lane that finds the items
// exhausted is done. Afterwards every lane is running a path or done.
template <class Work>
__device__ __forceinline__ void refill(PathState &P, const Tracer &T, const Work &work,
                                       uint32_t *counter) {
  const uint32_t lane = __lane_id();
  for (;;) {
    const bool need = P.stage == ST_IDLE;
    const uint64_t mask = __ballot(need);
    if (mask == 0) return;
    const uint32_t leader = (uint32_t)__ffsll((unsigned long long)mask) - 1u;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(mask));
    base = __builtin_amdgcn_readlane(base, leader);
    if (need) {
      const uint32_t rank = __builtin_amdgcn_mbcnt_hi(
          (uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
      const uint32_t item = base + rank;
      if (item < work.count()) {
        if (work.start(T, P, item)) P.item = item;
      } else {
        P.stage = ST_DONE;
      }
    }
  }
}

template <class Work, int W>
__global__ void __launch_bounds__(kBlock, W) k_paths(const lw_params p, const DevScene s,
                                                  const Work work,
                                                  uint32_t *__restrict__ counter) {
  __shared__ float lut_s[256];
  load_lut(lut_s, s.gamma_lut);
  const Tracer T{s, p, lut_s};
  PathState P;
  P.stage = ST_IDLE;
  P.item = 0;
  for (;;) {
    refill(P, T, work, counter);
    if (__ballot(P.stage != ST_DONE) == 0) break;
    if (P.stage == ST_EXTEND || P.stage == ST_SHADOW || P.stage == ST_EMITTER) {
      Hit h;
      const bool hit = trace(s, p, P.ro, P.rd, P.rtmin, P.rtmax, h, nullptr);
      T.consume(P, hit, h);
      if (P.stage == ST_IDLE) work.write(P.item, P.radiance);
    }
  }
}

// ---- the wavefront form: path state in HBM, one lean trace kernel per iteration ----
//
// Each of n slots holds a PathState. An iteration is three launches:
//   k_wf_logic  consumes each slot's last trace result (Tracer::consume), writes finished
//               paths' radiance, gives idle slots new work items (refill), and appends each
//               slot that now needs a ray to the ray queue -- compacted: the wave's ballot
//               counts its lanes with a ray, one atomic reserves their queue entries, each
//               lane writes at its rank;
//   k_wf_flip   makes the appended queue the one to trace;
//   k_wf_trace  traces the queued rays, one lane per ray, and stores each hit by slot.
// The trace kernel holds no path state, so it runs at the DDA's own register count and
// far more waves hide the latency of its dependent loads.
struct WfSlots {
  PathState *__restrict__ st;
  float4 *__restrict__ ray_o;  // o.xyz, tmin
  float4 *__restrict__ ray_d;  // d.xyz, tmax
  uint4 *__restrict__ hit;     // t bits, quad, texel, hit
  uint32_t *__restrict__ queue;
  uint32_t *__restrict__ ctr;  // [0] work items taken, [1] rays to trace, [2] rays appended
  uint32_t n;
};

__global__ void __launch_bounds__(kBlock) k_wf_init(WfSlots w) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < 4) w.ctr[i] = 0u;
  if (i >= w.n) return;
  w.st[i].stage = ST_IDLE;
}

template <class Work, int W>
__global__ void __launch_bounds__(kBlock, W) k_wf_logic(const lw_params p, const DevScene s,
                                                     const Work work, WfSlots w) {
  __shared__ float lut_s[256];
  load_lut(lut_s, s.gamma_lut);
  const Tracer T{s, p, lut_s};
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;  // n is a multiple of kBlock
  PathState P = w.st[i];
  if (P.stage == ST_DONE) return;
  if (P.stage == ST_EXTEND || P.stage == ST_SHADOW || P.stage == ST_EMITTER) {
    const uint4 r = w.hit[i];
    const Hit h{bits_f(r.x), r.y, r.z, 0.0f, 0.0f};
    T.consume(P, r.w != 0u, h);
    if (P.stage == ST_IDLE) work.write(P.item, P.radiance);
  }
  refill(P, T, work, &w.ctr[0]);
  w.st[i] = P;
  const bool pending = P.stage == ST_EXTEND || P.stage == ST_SHADOW || P.stage == ST_EMITTER;
  const uint64_t mask = __ballot(pending);
  if (mask == 0) return;
  const uint32_t lane = __lane_id();
  const uint32_t leader = (uint32_t)__ffsll((unsigned long long)mask) - 1u;
  uint32_t base = 0;
  if (lane == leader) base = atomicAdd(&w.ctr[2], (uint32_t)__popcll(mask));
  base = __builtin_amdgcn_readlane(base, leader);
  if (pending) {
    const uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                                    __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
    w.queue[base + rank] = i;
    w.ray_o[i] = make_float4(P.ro.x, P.ro.y, P.ro.z, P.rtmin);
    w.ray_d[i] = make_float4(P.rd.x, P.rd.y, P.rd.z, P.rtmax);
  }
}

__global__ void k_wf_flip(uint32_t *__restrict__ ctr) {
  if (threadIdx.x == 0) {
    ctr[1] = ctr[2];
    ctr[2] = 0u;
  }
}

template <int W>
__global__ void __launch_bounds__(kBlock, W) k_wf_trace(const lw_params p, const DevScene s,
                                                     WfSlots w) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= w.ctr[1]) return;
  const uint32_t i = w.queue[j];
  const float4 o = w.ray_o[i], d = w.ray_d[i];
  Hit h{INFINITY, UINT32_MAX, 0u, 0.0f, 0.0f};
  const bool hit = trace(s, p, V3{o.x, o.y, o.z}, V3{d.x, d.y, d.z}, o.w, d.w, h, nullptr);
  w.hit[i] = make_uint4(f_bits(h.t), h.quad, h.texel, hit ? 1u : 0u);
}

// ---- differentiable rendering (lw_diff.hpp) ----

// The forward under the parameters: k_render with render_sample_diff.
template <int W, int L>
__global__ void __launch_bounds__(kBlock, W) k_render_diff(const lw_params p, const DevScene s,
                                                        const DiffDev D, uint32_t sample_start,
                                                        float *__restrict__ rad) {
  __shared__ float lut_s[256];
  load_lut(lut_s, s.gamma_lut);
  const uint32_t w = p.frame[0], h = p.frame[1];
  const LaunchLane Ln = launch_lane<L>(w);
  const uint32_t px = Ln.px, py = Ln.py, sl = Ln.sl;
  if (px >= w || py >= h) return;
  const GradSink G{nullptr, 0u, 0u};
  const SuffixList sfx{nullptr, 0u};
  const V3 c = render_sample_diff<false>(s, p, p, lut_s, D, px, py, sample_start + sl, G, sfx,
                                         V3{0.0f, 0.0f, 0.0f});
  float *o = rad + ((size_t)sl * h * w + (size_t)py * w + px) * 3;
  o[0] = c.x;
  o[1] = c.y;
  o[2] = c.z;
}

// The adjoint: each lane replays one sample with dl = dL/dI(pixel) / spp; the workgroup
// sums its gradients in LDS (f64) and adds them to D.grads once per slot. Dynamic LDS:
// diff_slots(D) doubles.
template <int W, int L>
__global__ void __launch_bounds__(kBlock, W) k_adjoint(const lw_params p, const DevScene s,
                                                    const DiffDev D, uint32_t sample_start,
                                                    float inv_spp) {
  __shared__ float lut_s[256];
  __shared__ uint32_t sfx_s[kMaxSuffix * kBlock];
  extern __shared__ double grad_s[];
  const uint32_t ns = diff_slots(D);
  for (uint32_t i = threadIdx.x; i < ns; i += blockDim.x) grad_s[i] = 0.0;
  load_lut(lut_s, s.gamma_lut);  // ends with a barrier
  const uint32_t w = p.frame[0], h = p.frame[1];
  const LaunchLane Ln = launch_lane<L>(w);
  const uint32_t px = Ln.px, py = Ln.py;
  if (px < w && py < h) {
    const float *g = D.dldi + ((size_t)py * w + px) * 3;
    const V3 dl = V3{g[0], g[1], g[2]} * inv_spp;
    if (dl.x != 0.0f || dl.y != 0.0f || dl.z != 0.0f) {
      const GradSink G{grad_s, D.n_pal, D.n_int};
      const SuffixList sfx{sfx_s + threadIdx.x, (uint32_t)kBlock};
      render_sample_diff<true>(s, p, p, lut_s, D, px, py, sample_start + Ln.sl, G, sfx, dl);
    }
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < ns; i += blockDim.x) {
    const double v = grad_s[i];
    if (v != 0.0) atomicAdd(&D.grads[i], v);
  }
}

// The direct form with several samples per lane: lane (px, py) renders samples
// first .. first + count - 1 of its pixel one after another in one loop, starting the next
// sample as soon as its path ends, so a wave's lanes wait on one another's path lengths once
// per launch instead of once per sample. Each sample's arithmetic is render_sample's.
template <int W>
__global__ void __launch_bounds__(kBlock, W) k_render_multi(const lw_params p, const DevScene s,
                                                         uint32_t sample_start, uint32_t nb,
                                                         uint32_t per_lane,
                                                         float *__restrict__ rad) {
  __shared__ float lut_s[256];
  load_lut(lut_s, s.gamma_lut);
  const uint32_t w = p.frame[0], h = p.frame[1];
  const uint32_t tiles_x = (w + 15u) / 16u;
  const uint32_t tile = blockIdx.x;
  const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
  const uint32_t px = (tile % tiles_x) * 16u + (wave & 1u) * 8u + (lane & 7u);
  const uint32_t py = (tile / tiles_x) * 16u + (wave >> 1) * 8u + (lane >> 3);
  if (px >= w || py >= h) return;
  const uint32_t first = blockIdx.y * per_lane;
  const uint32_t count = min(per_lane, nb - first);
  SampleState S;
  uint32_t k = 0;
  bool live = false;
  for (;;) {
    if (!live) {
      if (k > 0) {
        float *o = rad + ((size_t)(first + k - 1) * h * w + (size_t)py * w + px) * 3;
        o[0] = S.radiance.x;
        o[1] = S.radiance.y;
        o[2] = S.radiance.z;
      }
      if (k == count) break;
      sample_begin(S, p, px, py, sample_start + first + k);
      k++;
    }
    live = sample_event(S, s, p, lut_s);
  }
}

// A batch pass's view, pixel and sample slice of workgroup `base`. The pass gives each view
// tiles x n workgroups (n = the view's samples this pass), workgroup q tiles + t standing for
// 16 x 16 tile t at sample q (k_render_batch's layout). When n is a multiple of 16, the 16
// workgroups 16 g .. 16 g + 15 of a tile are instead its sixteen 4 x 4 sub-tiles at samples
// 16 g .. 16 g + 15, with waves of 2 x 2 pixels x 16 samples (see launch_lane); with L 2, when
// every view's n is a multiple of 64, the 64 workgroups 64 g .. 64 g + 63 are its 2 x 2
// sub-tiles at samples 64 g .. 64 g + 63, a wave per pixel. The same (pixel, sample) pairs
// either way, so the staging and its accumulation are unchanged.
struct BatchItem {
  uint32_t v, sl, px, py;
  bool in_frame;
};
template <int L>
__device__ __forceinline__ BatchItem batch_item(const BatchDev *__restrict__ bd, uint32_t base) {
  BatchItem it;
  it.v = find_view(bd->item_prefix, bd->nviews, base);
  const uint32_t rest = (base - bd->item_prefix[it.v]) >> 8;
  const uint32_t tiles = bd->tiles[it.v], tx = bd->tiles_x[it.v];
  const uint32_t tile = rest % tiles, q = rest / tiles;
  const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63u;
  if constexpr (L == 2) {
    // every view's samples this pass a multiple of 64 (the host checks): workgroups
    // 64 g .. 64 g + 63 of a tile are its sixty-four 2 x 2 sub-tiles at samples 64 g .. + 63,
    // a wave per pixel
    const uint32_t sub = q & 63u;
    it.sl = (q & ~63u) + lane;
    it.px = (tile % tx) * 16u + (sub & 7u) * 2u + (wave & 1u);
    it.py = (tile / tx) * 16u + (sub >> 3) * 2u + (wave >> 1);
  } else if (bd->n[it.v] % 16u == 0) {
    const uint32_t sub = q & 15u, pix = lane & 3u;
    it.sl = (q & ~15u) + (lane >> 2);
    it.px = (tile % tx) * 16u + (sub & 3u) * 4u + (wave & 1u) * 2u + (pix & 1u);
    it.py = (tile / tx) * 16u + (sub >> 2) * 4u + (wave >> 1) * 2u + (pix >> 1);
  } else {
    it.sl = q;
    it.px = (tile % tx) * 16u + (wave & 1u) * 8u + (lane & 7u);
    it.py = (tile / tx) * 16u + (wave >> 1) * 8u + (lane >> 3);
  }
  it.in_frame = it.px < bd->cam[it.v].width && it.py < bd->cam[it.v].height;
  return it;
}

// The differentiable forward of a batch pass (k_render_batch with render_sample_diff).
template <int W, int L>
__global__ void __launch_bounds__(kBlock, W) k_render_diff_batch(const lw_params p, const DevScene s,
                                                              const DiffDev D,
                                                              const BatchDev *__restrict__ bd,
                                                              float *__restrict__ rad) {
  const uint32_t base = blockIdx.x * (uint32_t)kBlock;
  if (base >= bd->total) return;
  __shared__ float lut_s[256];
  load_lut(lut_s, s.gamma_lut);
  const BatchItem it = batch_item<L>(bd, base);
  if (!it.in_frame) return;
  const ViewCam cam = bd->cam[it.v];
  const GradSink G{nullptr, 0u, 0u};
  const SuffixList sfx{nullptr, 0u};
  const V3 c = render_sample_diff<false>(s, p, cam, lut_s, D, it.px, it.py, bd->start[it.v] + it.sl,
                                         G, sfx, V3{0.0f, 0.0f, 0.0f});
  float *o = rad + bd->rad_base[it.v] +
             ((size_t)it.sl * cam.height * cam.width + (size_t)it.py * cam.width + it.px) * 3;
  o[0] = c.x;
  o[1] = c.y;
  o[2] = c.z;
}

// The adjoint of a batch pass: D.dldi holds every view's dL/dI, view v's pixels from
// pix_prefix[v]; each sample's dl is its pixel's dL/dI / the view's spp.
template <int W, int L>
__global__ void __launch_bounds__(kBlock, W) k_adjoint_batch(const lw_params p, const DevScene s,
                                                          const DiffDev D,
                                                          const BatchDev *__restrict__ bd) {
  const uint32_t base = blockIdx.x * (uint32_t)kBlock;
  if (base >= bd->total) return;  // uniform over the workgroup: no barrier is skipped by part
  __shared__ float lut_s[256];
  __shared__ uint32_t sfx_s[kMaxSuffix * kBlock];
  extern __shared__ double grad_s[];
  const uint32_t ns = diff_slots(D);
  for (uint32_t i = threadIdx.x; i < ns; i += blockDim.x) grad_s[i] = 0.0;
  load_lut(lut_s, s.gamma_lut);  // ends with a barrier
  const BatchItem it = batch_item<L>(bd, base);
  if (it.in_frame) {
    const ViewCam cam = bd->cam[it.v];
    const float *g = D.dldi + ((size_t)bd->pix_prefix[it.v] + (size_t)it.py * cam.width + it.px) * 3;
    const V3 dl = V3{g[0], g[1], g[2]} * (1.0f / (float)cam.spp);
    if (dl.x != 0.0f || dl.y != 0.0f || dl.z != 0.0f) {
      const GradSink G{grad_s, D.n_pal, D.n_int};
      const SuffixList sfx{sfx_s + threadIdx.x, (uint32_t)kBlock};
      render_sample_diff<true>(s, p, cam, lut_s, D, it.px, it.py, bd->start[it.v] + it.sl, G, sfx, dl);
    }
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < ns; i += blockDim.x) {
    const double v = grad_s[i];
    if (v != 0.0) atomicAdd(&D.grads[i], v);
  }
}

// The L2 loss of a batch against its targets, L = sum over views, pixels and channels of
// (I - T)^2 / N with I = the f64 sum / spp: dL/dI into dldi (f32) and L into *loss (f64,
// one atomic per workgroup).
__global__ void __launch_bounds__(kBlock) k_fit_loss(const BatchDev *__restrict__ bd,
                                                     const double *__restrict__ acc,
                                                     const float *__restrict__ target,
                                                     float *__restrict__ dldi, double inv_n,
                                                     double *__restrict__ loss) {
  __shared__ double part[kBlock / 64];
  const uint32_t g3 = blockIdx.x * blockDim.x + threadIdx.x;
  double sq = 0.0;
  if (g3 < 3u * bd->npix_total) {
    const uint32_t v = find_view(bd->pix_prefix, bd->nviews, g3 / 3u);
    const double d = acc[g3] / (double)bd->cam[v].spp - (double)target[g3];
    dldi[g3] = (float)(2.0 * d * inv_n);
    sq = d * d * inv_n;
  }
  for (int off = 32; off > 0; off >>= 1) sq += __shfl_down(sq, off);
  if ((threadIdx.x & 63u) == 0) part[threadIdx.x >> 6] = sq;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < kBlock / 64; w++) t += part[w];
    atomicAdd(loss, t);
  }
}

// One Adam step (D. P. Kingma, J. Ba, "Adam: A Method for Stochastic Optimization", ICLR
// 2015) on the factors x (f32) from their gradient g (f64), with moments m, v; x stays at
// or above lo. bc1, bc2 = 1 - beta^t.
__global__ void __launch_bounds__(kBlock) k_adam(float *__restrict__ x, const double *__restrict__ g,
                                                 float *__restrict__ m, float *__restrict__ v,
                                                 uint32_t n, float lr, float beta1, float beta2,
                                                 float eps, float bc1, float bc2, float lo) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = (float)g[i];
  const float mi = beta1 * m[i] + (1.0f - beta1) * gi;
  const float vi = beta2 * v[i] + (1.0f - beta2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float step = lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
  x[i] = fmaxf(x[i] - step, lo);
}

}  // namespace

// ------------------------------------------------------------------------------------
// host side

namespace {

thread_local std::string g_last_error;

struct HipError {
  hipError_t code;
  std::string what;
};

#define HIP_TRY(expr)                                                                  \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess) throw HipError{e_, std::string(#expr) + ": " + hipGetErrorString(e_)}; \
  } while (0)

struct InvalidScene {
  std::string what;
};
struct Unsupported {  // a valid request this backend does not serve (LW_EUNSUPPORTED)
  std::string what;
};

template <typename T>
struct DevBuf {
  T *ptr = nullptr;
  size_t n = 0;
  void alloc(size_t count) {
    free();
    n = count;
    if (count) {
      hipError_t e = hipMalloc((void **)&ptr, count * sizeof(T));
      if (e != hipSuccess) {
        ptr = nullptr;
        n = 0;
        throw HipError{e, "hipMalloc(" + std::to_string(count * sizeof(T)) + " bytes)"};
      }
    }
  }
  void upload(const T *src, size_t count, hipStream_t st) {
    alloc(count ? count : 1);
    if (count) HIP_TRY(hipMemcpyAsync(ptr, src, count * sizeof(T), hipMemcpyHostToDevice, st));
  }
  void ensure(size_t count) {
    if (count > n) alloc(count);
  }
  void adopt(T *p, size_t count) {  // takes ownership of a hipMalloc'ed buffer
    free();
    ptr = p;
    n = count;
  }
  void free() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    n = 0;
  }
  ~DevBuf() { free(); }
};

// The params fields that decide which buffer words a kernel reads; a resident scene is
// validated once for these, so a render's params must repeat them.
struct LayoutKey {
  uint32_t grid_dims[4], off[4], off2[4], see_through[4], strategy_on;
  float sky_res, sky_map[4];
  bool operator==(const LayoutKey &o) const { return std::memcmp(this, &o, sizeof(*this)) == 0; }
};

LayoutKey layout_key(const lw_params &p) {
  LayoutKey k;
  std::memset(&k, 0, sizeof(k));
  std::memcpy(k.grid_dims, p.grid_dims, sizeof(k.grid_dims));
  std::memcpy(k.off, p.off, sizeof(k.off));
  std::memcpy(k.off2, p.off2, sizeof(k.off2));
  std::memcpy(k.see_through, p.see_through, sizeof(k.see_through));
  k.strategy_on = (p.misc_u[3] > 0 && p.emitter[0] != LW_EMITTER_SAMPLING_NONE) ? 1u : 0u;
  k.sky_res = p.misc_f[2];
  std::memcpy(k.sky_map, p.sky_map, sizeof(k.sky_map));
  return k;
}

}  // namespace

struct lw_hip_scene {
  DevBuf<uint32_t> grid;
  DevBuf<lw_model> models;
  DevBuf<lw_quad> quads;
  DevBuf<uint32_t> qinfo;
  DevBuf<uint32_t> atlas;
  DevBuf<lw_emitter> emitters;
  DevBuf<uint32_t> egrid;
  DevBuf<float> sky;
  DevBuf<float> lut;
  DevBuf<uint2> occupancy;
  DevBuf<uint8_t> octant_boxes;
  DevBuf<uint32_t> mfaces;
  DevBuf<uint32_t> cell_ext;
  std::vector<uint32_t> quad_flags;  // host copy of each quad's flags (p0.w)
  // The params layout the scene was validated with. A scene built on the device
  // (lw_hip_scene_build) takes it from the first render's params, checked against `built`.
  mutable LayoutKey key;
  mutable bool key_set = false;
  bool is_built = false;
  lw_scene_info built{};
  size_t atlas_len = 0, sky_len = 0;
  uint32_t max_emitter_rays = 1;  // emitter face rays one bounce can trace
  DevScene dev() const {
    return DevScene{grid.ptr, models.ptr, reinterpret_cast<const float4 *>(quads.ptr), qinfo.ptr,
                    atlas.ptr, emitters.ptr, egrid.ptr, sky.ptr, lut.ptr, occupancy.ptr, octant_boxes.ptr,
                    mfaces.ptr, cell_ext.ptr};
  }
};

// A batch's views laid out for the device and its passes: as few as fit the staging cap,
// each with fewer than 2^31 work items.
struct BatchPlan {
  std::vector<BatchDev> host = std::vector<BatchDev>(1);
  uint32_t npass = 0;
  uint64_t items = 0, staging = 0, npix_total = 0, samples = 0;
  BatchDev &bd() { return host[0]; }
  const BatchDev &bd() const { return host[0]; }
};

struct lw_hip_backend {
  int device = 0;
  std::string name;
  std::string last_error;
  hipStream_t stream = nullptr;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  DevBuf<float> rad;
  DevBuf<double> acc;
  DevBuf<float> out;
  DevBuf<uint32_t> counter;
  lw_timing timing{};
  // Launch shape: the persistent path kernels run `blocks_*` workgroups of kBlock lanes
  // (the device's resident capacity).
  uint32_t blocks_frame = 0, blocks_list = 0;
  // LW_KERNEL: "direct" (one lane per sample, the default and the fastest measured at the
  // 720p 256-spp frame), "persistent" (k_paths) or "wavefront" (k_wf_*, LW_WF_SLOTS path
  // slots).
  int mode = 1;
  // LW_WAVES, LW_TRACE_WAVES: register budgets (see dispatch). The direct form measured
  // fastest at 6 waves per SIMD (80 VGPRs): 564 Msamples/s at the 720p 256-spp frame
  // against 528 at 5, 516 at 7 and 412 at 8. The persistent form, which has no 6-wave variant, then
  // runs at its own best, 4.
  int waves = 6, trace_waves = 1;
  // LW_TILES_WAVES (or LW_WAVES): the persistent tile kernel's budget, 4 to 6 waves per SIMD
  // (other values are brought into that range). 5 measured fastest: 1146 Msamples/s at the
  // 720p 256-spp frame with one-pixel tiles against 1094 at 4 and 1088 at 6 (the launch ramp's
  // k_render stays fastest at 6).
  int tiles_waves = 5;
  // LW_XCD_REMAP=1: the direct kernel gives each XCD a contiguous run of tiles.
  uint32_t xcd_remap = 0;
  // LW_SPL: samples per lane of the direct form (k_render_multi when above 1). Measured
  // slower at every budget (at 6 waves: 564, 509, 495, 475 Msamples/s for 1, 4, 8, 16): lanes
  // whose paths drift apart in depth lose more to divergence than the wave gains back.
  uint32_t per_lane = 1;
  uint32_t wf_slots = 1u << 20;
  DevBuf<PathState> wf_st;
  DevBuf<float4> wf_o, wf_d;
  DevBuf<uint4> wf_hit;
  DevBuf<uint32_t> wf_queue, wf_ctr;
  uint32_t *pinned = nullptr;
  void *scratch = nullptr;  // this device's scratch array (hipGetSymbolAddress), or null
  uint64_t last_iterations = 0;
  // Batches: the staging cap, a minimum pass count (0: as few as fit) and whether a pass
  // replays as a captured hipGraph (LW_BATCH_GRAPHS=0 launches it kernel by kernel). 8 GiB of
  // the 288 GB HBM lets the cfg4 fit (32 views of 640 x 360 at 128 spp, 88.5 MB of staging
  // per sample per pixel) run in 2 passes of 64 samples, a wave per pixel, instead of 4 of 32.
  uint64_t batch_staging = 8ull << 30;
  uint32_t batch_passes = 0;
  bool use_graphs = true;
  // LW_BATCH_FORM=passes: the direct form's batches in passes (k_render_batch) rather than in
  // one persistent launch over (view, tile) items (k_render_tiles).
  bool batch_passes_form = false;
  // LW_FRAME_FORM=launches: frames in launches of up to 64 samples per pixel with staging and
  // an f64 accumulation pass (render/cuda.rs's ramp) rather than in one persistent
  // k_render_tiles launch.
  bool frame_launches = false;
  uint32_t *progress_host = nullptr;  // host memory k_render_tiles reports its items to
  uint32_t *progress_dev = nullptr;   // the same memory's device address
  DevBuf<uint32_t> batch_items;
  uint32_t blocks_tiles = 0;  // resident workgroups of k_render_tiles
  DevBuf<BatchDev> batch_dev;
  DevBuf<float> tone_in;
  DevBuf<uint32_t> tone_out;
  // libm's corrections to the rounded f64 value per function (POW_ENCODE, POW_DECODE,
  // LIBM_LOG2; see lw_libm.hpp), made on the first call that needs them.
  struct PowFix {
    DevBuf<uint2> fix;
    uint32_t n = 0;
    bool ready = false;
    double ms = 0.0;
  } pow_fix[3];
  BatchPlan plan;  // the last batch's layout (its sums stay in acc when it keeps them)
  // Differentiable rendering: the parameters and derivative inputs of the last call.
  DevBuf<float> diff_pal, diff_inten, diff_dsky, diff_dldi;
  DevBuf<uint32_t> diff_qg, diff_qig;
  DevBuf<double> diff_grads;
  // Quick mode: the frame's inputs and output, and the rasterizer's own buffers (grow-only).
  DevBuf<lw_quick_vertex> quick_verts;
  DevBuf<uint32_t> quick_idx, quick_atlas, quick_mips, quick_stars, quick_out;
  lwq::Workspace *quick_ws = nullptr;
};

namespace {

void set_err(lw_hip_backend *b, const std::string &msg) {
  if (b) b->last_error = msg;
  g_last_error = msg;
}

// Whether every scene buffer is smaller than 4 GiB (element counts of the grid, models, quads,
// atlas, emitters, emitter grid and sky), as the kernels' 32-bit byte offsets need.
bool buffers_fit_32bit(uint64_t grid, uint64_t models, uint64_t quads, uint64_t atlas,
                       uint64_t emitters, uint64_t egrid, uint64_t sky) {
  const uint64_t lim = 0xFFFFFFFFull;
  return grid * 4 <= lim && models * sizeof(lw_model) <= lim && quads * sizeof(lw_quad) <= lim &&
         atlas * 4 <= lim && emitters * sizeof(lw_emitter) <= lim && egrid * 4 <= lim && sky * 4 <= lim;
}

// Every index the kernels compute from the scene buffers stays in bounds: the checks the
// reference's slice indexing makes at run time (a panic there), made once here.
void validate(const lw_scene_view &v) {
  auto fail = [](const std::string &m) { throw InvalidScene{m}; };
  if (!v.params) fail("params is null");
  const lw_params &p = *v.params;
  const uint64_t glen = v.grid_len;
  if (!v.grid || !v.models || !v.quads || !v.atlas || !v.emitters || !v.emitter_grid)
    fail("a scene buffer is null");
  if (v.models_len == 0 || v.quads_len == 0 || v.emitters_len == 0 || v.atlas_len == 0)
    fail("models, quads, emitters and atlas must not be empty");
  // The kernels address every scene buffer by a 32-bit byte offset from its base.
  if (!buffers_fit_32bit(v.grid_len, v.models_len, v.quads_len, v.atlas_len, v.emitters_len,
                         v.emitter_grid_len, v.sky ? v.sky_len : 0))
    throw Unsupported{"a scene buffer of 4 GiB or more"};
  const uint64_t dx = p.grid_dims[0], dy = p.grid_dims[1], dz = p.grid_dims[2];
  if (dx == 0 || dy == 0 || dz == 0 || dx * 16 > 0x7FFFFFF || dy * 16 > 0x7FFFFFF ||
      dz * 16 > 0x7FFFFFF)
    fail("grid_dims out of range");
  const uint64_t nb = dx * dy * dz;
  if (nb > 0xFFFFFFFFull) fail("too many bricks");
  if ((uint64_t)p.off[0] + nb > glen || (uint64_t)p.off[2] + nb > glen)
    fail("top-level or overlay entries out of the grid buffer");
  // models: every quad of every model is in the quad buffer.
  for (size_t m = 0; m < v.models_len; m++)
    if ((uint64_t)v.models[m].quad_offset + v.models[m].quad_count > v.quads_len)
      fail("model " + std::to_string(m) + " names quads past the quad buffer");
  for (uint64_t b = 0; b < nb; b++) {
    const uint32_t e = v.grid[p.off[0] + b];
    if (e != LW_EMPTY) {
      if (e & LW_UNIFORM) {
        if ((e & ~LW_UNIFORM) >= v.models_len) fail("uniform brick with an unknown model");
      } else {
        const uint64_t base = (uint64_t)p.off[1] + (uint64_t)e * LW_BRICK_WORDS;
        if (base + LW_BRICK_WORDS > glen || base + LW_BRICK_WORDS > 0xFFFFFFFFull)
          fail("brick index past the grid buffer");
        for (uint32_t k = 0; k < LW_BRICK_WORDS; k++) {
          const uint32_t w = v.grid[base + k];
          if ((w & 0xFFFFu) >= v.models_len || (w >> 16) >= v.models_len)
            fail("voxel with an unknown model id");
        }
      }
    }
    const uint32_t o = v.grid[p.off[2] + b];
    if (o != LW_EMPTY) {
      const uint64_t base = (uint64_t)p.off[3] + (uint64_t)o * LW_OV_CELL_WORDS;
      if (base + LW_OV_CELL_WORDS > glen || base + LW_OV_CELL_WORDS > 0xFFFFFFFFull)
        fail("overlay brick past the grid buffer");
      for (uint32_t c = 0; c < 4096; c++) {
        const uint64_t off = v.grid[base + 2 * c], cnt = v.grid[base + 2 * c + 1];
        const uint64_t first = (uint64_t)p.off2[0] + off;
        if (first + cnt > glen || first + cnt > 0xFFFFFFFFull) fail("overlay cell past the grid buffer");
        for (uint64_t i = 0; i < cnt; i++)
          if (v.grid[first + i] >= v.quads_len) fail("overlay quad index past the quad buffer");
      }
    }
  }
  const int64_t aw = (int32_t)p.off2[1], ah = (int32_t)p.off2[2];
  if (aw < 1 || ah < 1 || (uint64_t)(aw * ah) > v.atlas_len) fail("atlas size does not fit the atlas");
  if ((uint64_t)p.see_through[0] + p.see_through[1] > v.quads_len) fail("see-through range past the quads");
  // emitters: the grid's cells, lists and emitters stay in their buffers.
  if (p.misc_u[3] > 0 && p.emitter[0] != LW_EMITTER_SAMPLING_NONE) {
    const uint32_t *g = v.emitter_grid;
    if (v.emitter_grid_len < LW_EMITTER_GRID_HEADER) fail("emitter grid shorter than its header");
    const uint64_t cells = (uint64_t)g[1] * g[2] * g[3];
    const uint64_t list = LW_EMITTER_GRID_HEADER + cells * 2;
    if (list > v.emitter_grid_len) fail("emitter grid cells past the buffer");
    if ((int32_t)g[1] < 0 || (int32_t)g[2] < 0 || (int32_t)g[3] < 0) fail("emitter grid size");
    for (uint64_t c = 0; c < cells; c++) {
      const uint64_t start = g[LW_EMITTER_GRID_HEADER + 2 * c], cnt = g[LW_EMITTER_GRID_HEADER + 2 * c + 1];
      if (list + start + cnt > v.emitter_grid_len) fail("emitter cell list past the buffer");
      for (uint64_t k = 0; k < cnt; k++)
        if (g[list + start + k] >= v.emitters_len) fail("emitter index past the emitter buffer");
    }
    for (size_t e = 0; e < v.emitters_len; e++) {
      uint32_t first;
      std::memcpy(&first, &v.emitters[e].pos[3], 4);
      if ((uint64_t)first + v.emitters[e].faces[0] > v.quads_len) fail("emitter faces past the quads");
    }
  }
  // sky: the cache or the two maps the integrator reads.
  if (p.sky_map[0] >= 1.0f) {
    for (int k = 1; k < 4; k++)
      if (!(p.sky_map[k] >= 1.0f)) fail("sky map sizes must be at least 1");
    const uint64_t w0 = as_u32(p.sky_map[0]), h0 = as_u32(p.sky_map[1]);
    const uint64_t w1 = as_u32(p.sky_map[2]), h1 = as_u32(p.sky_map[3]);
    const uint64_t base = (uint64_t)as_u32(p.sky_map[0] * p.sky_map[1]) * 3;
    if (w0 * h0 * 3 > v.sky_len || base + w1 * h1 * 3 > v.sky_len || !v.sky) fail("sky maps past the sky buffer");
  } else if (p.misc_f[2] >= 1.0f) {
    const uint64_t m = (uint64_t)as_u32(p.misc_f[2]) + 1;
    if (m * m * 3 > v.sky_len || !v.sky) fail("sky cache past the sky buffer");
  }
}

// The axis code of a quad (QI_* in lw_device.hpp).
uint32_t quad_info(const lw_quad &q) {
  for (int c = 0; c < 4; c++)
    for (const float *a : {q.p0, q.eu, q.ev})
      if (!std::isfinite(a[c]) && c < 3) return 0u;
  auto axis_of = [](const float *e, int &axis, bool &neg) {
    int nz = 0;
    for (int c = 0; c < 3; c++) {
      if (e[c] == 1.0f || e[c] == -1.0f) {
        axis = c;
        neg = e[c] < 0.0f;
        nz++;
      } else if (e[c] != 0.0f) {
        return false;
      }
    }
    return nz == 1;
  };
  int i, j;
  bool nu, nv;
  if (!axis_of(q.eu, i, nu) || !axis_of(q.ev, j, nv) || i == j) return 0u;
  const int k = 3 - i - j;
  // eu x ev = su sv (e_i x e_j) = su sv eps_ijk e_k
  const bool cyclic = (j == (i + 1) % 3);
  const bool nn = (nu != nv) != !cyclic;
  return QI_AXIS | (uint32_t)k | ((uint32_t)i << 2) | ((uint32_t)j << 4) | (nu ? QI_NEG_U : 0u) |
         (nv ? QI_NEG_V : 0u) | (nn ? QI_NEG_N : 0u);
}

// Each model's face word (DevScene::mfaces, MF_* in lw_device.hpp) from its quads' axis codes
// and flags: set for a model whose quads are all FRONT_ONLY axis quads with distinct normals
// (every full cube), 0 for any other.
std::vector<uint32_t> model_faces(const lw_model *models, size_t n, const std::vector<uint32_t> &info,
                                  const std::vector<uint32_t> &flags) {
  std::vector<uint32_t> out(std::max<size_t>(n, 1), 0u);
  for (size_t m = 0; m < n; m++) {
    const lw_model &md = models[m];
    if (md.quad_count == 0 || md.quad_count > 6) continue;
    uint32_t fw = 0x3FFFFu;  // every direction MF_NONE
    bool ok = true;
    for (uint32_t i = 0; i < md.quad_count && ok; i++) {
      const uint64_t q = (uint64_t)md.quad_offset + i;
      if (q >= info.size() || !(info[q] & QI_AXIS) || !(flags[q] & LW_FLAG_FRONT_ONLY)) {
        ok = false;
        break;
      }
      const uint32_t f = 3u * (2u * (info[q] & 3u) + ((info[q] & QI_NEG_N) ? 1u : 0u));
      if (((fw >> f) & 7u) != MF_NONE) ok = false;
      fw = (fw & ~(7u << f)) | (i << f);
    }
    if (ok) out[m] = fw | MF_FACES;
  }
  return out;
}

// The number of stored bricks a view's top-level entries name (one past the largest entry).
uint64_t stored_bricks(const lw_scene_view &v) {
  const lw_params &p = *v.params;
  const uint64_t nb = (uint64_t)p.grid_dims[0] * p.grid_dims[1] * p.grid_dims[2];
  uint64_t stored = 0;
  for (uint64_t b = 0; b < nb; b++) {
    const uint32_t e = v.grid[p.off[0] + b];
    if (e != LW_EMPTY && (e & LW_UNIFORM) == 0) stored = std::max<uint64_t>(stored, (uint64_t)e + 1);
  }
  return stored;
}

// The sub-block occupancy of every stored brick (DevScene::occupancy), indexed by the
// brick's top-level entry.
std::vector<uint2> occupancy_masks(const lw_scene_view &v) {
  const lw_params &p = *v.params;
  const uint64_t stored = stored_bricks(v);
  std::vector<uint2> occ(std::max<uint64_t>(stored, 1), make_uint2(0u, 0u));
  for (uint64_t e = 0; e < stored; e++) {
    const uint32_t *words = v.grid + p.off[1] + e * LW_BRICK_WORDS;
    uint64_t m = 0;
    for (uint32_t local = 0; local < 4096; local++) {
      const uint32_t id = (words[local >> 1] >> ((local & 1u) * 16u)) & 0xFFFFu;
      if (id != 0) {
        const uint32_t x = local & 15u, z = (local >> 4) & 15u, y = local >> 8;
        m |= 1ull << (((y >> 2) * 4u + (z >> 2)) * 4u + (x >> 2));
      }
    }
    occ[e] = make_uint2((uint32_t)m, (uint32_t)(m >> 32));
  }
  return occ;
}

// DevScene::cell_ext of a scene whose grid is on the device, for its `stored` bricks; left null
// (the walk then crosses such cells one by one) when the table would not fit 32-bit offsets.
void build_cell_ext(lw_hip_backend *b, lw_hip_scene &sc, uint64_t off_bricks, uint64_t stored) {
  sc.cell_ext.free();
  if (stored == 0 || stored * 4096u * sizeof(uint32_t) >= (1ull << 32)) return;
  sc.cell_ext.alloc(stored * 4096u);
  k_cell_ext<<<dim3((unsigned)stored), dim3(kBlock), 0, b->stream>>>(sc.grid.ptr, off_bricks, sc.cell_ext.ptr);
  HIP_TRY(hipGetLastError());
}

// DevScene::octant_boxes of a scene whose grid is on the device.
void build_octant_boxes(lw_hip_backend *b, lw_hip_scene &sc, const uint32_t dims[3], uint32_t off_top,
                        uint32_t off_ov) {
  const uint64_t nb = (uint64_t)dims[0] * dims[1] * dims[2];
  sc.octant_boxes.alloc(std::max<uint64_t>(nb, 1));
  const auto blocks = [](uint64_t n) {
    return dim3((unsigned)std::min<uint64_t>(std::max<uint64_t>((n + kBlock - 1) / kBlock, 1), 1u << 20));
  };
  k_octants_init<<<blocks(nb), dim3(kBlock), 0, b->stream>>>(sc.grid.ptr, off_top, off_ov, nb, sc.octant_boxes.ptr);
  HIP_TRY(hipGetLastError());
  const uint64_t lines[3] = {(uint64_t)dims[1] * dims[2], (uint64_t)dims[2] * dims[0], (uint64_t)dims[1] * dims[0]};
  for (uint32_t axis = 0; axis < 3; axis++) {
    k_octants_scan<<<blocks(lines[axis]), dim3(kBlock), 0, b->stream>>>(sc.octant_boxes.ptr, dims[0], dims[1],
                                                                         dims[2], axis);
    HIP_TRY(hipGetLastError());
  }
  HIP_TRY(hipStreamSynchronize(b->stream));  // a scene is whole when its call returns
}

void upload_scene(lw_hip_backend *b, const lw_scene_view &v, lw_hip_scene &sc) {
  validate(v);
  HIP_TRY(hipSetDevice(b->device));
  sc.grid.upload(v.grid, v.grid_len, b->stream);
  sc.models.upload(v.models, v.models_len, b->stream);
  sc.quads.upload(v.quads, v.quads_len, b->stream);
  std::vector<uint32_t> info(v.quads_len);
  for (size_t q = 0; q < v.quads_len; q++) info[q] = quad_info(v.quads[q]);
  sc.quad_flags.resize(v.quads_len);
  for (size_t q = 0; q < v.quads_len; q++) std::memcpy(&sc.quad_flags[q], &v.quads[q].p0[3], 4);
  sc.qinfo.upload(info.data(), info.size(), b->stream);
  const std::vector<uint32_t> faces = model_faces(v.models, v.models_len, info, sc.quad_flags);
  sc.mfaces.upload(faces.data(), faces.size(), b->stream);
  sc.atlas.upload(v.atlas, v.atlas_len, b->stream);
  sc.emitters.upload(v.emitters, v.emitters_len, b->stream);
  sc.egrid.upload(v.emitter_grid, v.emitter_grid_len, b->stream);
  sc.sky.upload(v.sky, v.sky ? v.sky_len : 0, b->stream);
  float lut[256];
  gamma_lut(lut);
  sc.lut.upload(lut, 256, b->stream);
  const std::vector<uint2> occ = occupancy_masks(v);
  sc.occupancy.upload(occ.data(), occ.size(), b->stream);
  build_cell_ext(b, sc, v.params->off[1], stored_bricks(v));
  HIP_TRY(hipStreamSynchronize(b->stream));  // the host vectors above die with this scope
  build_octant_boxes(b, sc, v.params->grid_dims, v.params->off[0], v.params->off[2]);
  sc.key = layout_key(*v.params);
  sc.key_set = true;
  // The most face rays one bounce's emitter sampling can trace (integrator.rs:441-469).
  uint64_t max_faces = 0, max_count = 0;
  for (size_t e = 0; e < v.emitters_len; e++) max_faces = std::max<uint64_t>(max_faces, v.emitters[e].faces[0]);
  if (v.emitter_grid_len >= LW_EMITTER_GRID_HEADER) {
    const uint32_t *g = v.emitter_grid;
    const uint64_t cells = (uint64_t)g[1] * g[2] * g[3];
    for (uint64_t c = 0; c < cells && LW_EMITTER_GRID_HEADER + 2 * c + 1 < v.emitter_grid_len; c++)
      max_count = std::max<uint64_t>(max_count, g[LW_EMITTER_GRID_HEADER + 2 * c + 1]);
  }
  const uint64_t rays = max_faces * max_count;  // ALL's bound covers ONE and ONE_BLOCK
  sc.max_emitter_rays = (uint32_t)std::min<uint64_t>(std::max<uint64_t>(rays, 1), 1u << 20);
}


// Params must repeat the layout a scene was validated with. A built scene is validated
// against the first params it meets: the scene fields must be those the build reported, and
// the params-only checks of validate() (atlas size, sky layout) are made here.
void check_layout(const lw_hip_scene &sc, const lw_params &p) {
  if (!sc.key_set) {
    auto fail = [](const std::string &m) { throw InvalidScene{m}; };
    const lw_scene_info &i = sc.built;
    for (int a = 0; a < 3; a++)
      if (p.grid_dims[a] != i.dims[a]) fail("params grid_dims differ from the built scene");
    if (p.grid_dims[3] != i.dims[0] * i.dims[1] * i.dims[2]) fail("params grid_dims[3] is not the brick count");
    for (int k = 0; k < 4; k++)
      if (p.off[k] != i.offsets[k]) fail("params offsets differ from the built scene");
    if (p.off2[0] != i.offsets[4] || p.off2[3] != i.quads_len) fail("params offsets differ from the built scene");
    if (p.see_through[0] != i.see_through[0] || p.see_through[1] != i.see_through[1])
      fail("params see-through range differs from the built scene");
    if (p.misc_u[2] != i.models_len || p.misc_u[3] != i.emitter_count)
      fail("params model or emitter count differs from the built scene");
    const int64_t aw = (int32_t)p.off2[1], ah = (int32_t)p.off2[2];
    if (aw < 1 || ah < 1 || (uint64_t)(aw * ah) > sc.atlas_len) fail("atlas size does not fit the atlas");
    if (p.sky_map[0] >= 1.0f) {
      for (int k = 1; k < 4; k++)
        if (!(p.sky_map[k] >= 1.0f)) fail("sky map sizes must be at least 1");
      const uint64_t w0 = as_u32(p.sky_map[0]), h0 = as_u32(p.sky_map[1]);
      const uint64_t w1 = as_u32(p.sky_map[2]), h1 = as_u32(p.sky_map[3]);
      const uint64_t base = (uint64_t)as_u32(p.sky_map[0] * p.sky_map[1]) * 3;
      if (w0 * h0 * 3 > sc.sky_len || base + w1 * h1 * 3 > sc.sky_len) fail("sky maps past the sky buffer");
    } else if (p.misc_f[2] >= 1.0f) {
      const uint64_t m = (uint64_t)as_u32(p.misc_f[2]) + 1;
      if (m * m * 3 > sc.sky_len) fail("sky cache past the sky buffer");
    }
    sc.key = layout_key(p);
    sc.key_set = true;
  }
  if (!(layout_key(p) == sc.key))
    throw InvalidScene{"params do not match the layout the scene was validated with"};
}

// lw_hip_scene_build: the inputs checked as validate() checks an uploaded view, the volume
// and the models' buffers uploaded, the layout made on the device (lw_build.hip).
void build_scene(lw_hip_backend *b, const lw_voxel_scene &v, lw_hip_scene &sc, lw_scene_info &info,
                 lw_build_stats *stats) {
  const auto t0 = std::chrono::steady_clock::now();
  auto fail = [](const std::string &m) { throw InvalidScene{m}; };
  const uint64_t dx = v.dims[0], dy = v.dims[1], dz = v.dims[2], nb = dx * dy * dz;
  if (dx == 0 || dy == 0 || dz == 0 || dx * 16 > 0x7FFFFFF || dy * 16 > 0x7FFFFFF || dz * 16 > 0x7FFFFFF)
    fail("dims out of range");
  if (nb >= (1ull << 24)) fail("a built scene has fewer than 2^24 bricks");
  if (!v.voxels || !v.models || !v.model_flags || !v.atlas) fail("a scene buffer is null");
  if (v.models_len == 0 || v.models_len > 65536) fail("models_len out of range (1 .. 65536)");
  if (v.model_flags[0] != 0) fail("model 0 is air: its flags must be 0");
  if ((v.quads_len && !v.quads) || (v.entity_quads_len && !v.entity_quads)) fail("a quad buffer is null");
  if (v.see_through_count > v.entity_quads_len) fail("more see-through quads than entity quads");
  const uint64_t nq = (uint64_t)v.quads_len + v.entity_quads_len;
  if (nq >= 0xFFFFFFFFull) fail("too many quads");
  for (size_t m = 0; m < v.models_len; m++)
    if ((uint64_t)v.models[m].quad_offset + v.models[m].quad_count > v.quads_len)
      fail("model " + std::to_string(m) + " names quads past the quad buffer");
  if (v.atlas_len == 0 || v.atlas_size[0] == 0 || v.atlas_size[1] == 0 ||
      (uint64_t)v.atlas_size[0] * v.atlas_size[1] > v.atlas_len)
    fail("atlas size does not fit the atlas");
  HIP_TRY(hipSetDevice(b->device));
  DevBuf<uint16_t> vox;
  DevBuf<uint8_t> occ, flags;
  vox.upload(v.voxels, nb * 4096, b->stream);
  if (v.brick_occludes) occ.upload(v.brick_occludes, nb, b->stream);
  flags.upload(v.model_flags, v.models_len, b->stream);
  sc.models.upload(v.models, v.models_len, b->stream);
  std::vector<lw_quad> quads(v.quads, v.quads + v.quads_len);
  quads.insert(quads.end(), v.entity_quads, v.entity_quads + v.entity_quads_len);
  if (quads.empty()) quads.push_back(lw_quad{});  // voxel.rs:441-443
  sc.quads.upload(quads.data(), quads.size(), b->stream);
  std::vector<uint32_t> qi(quads.size());
  sc.quad_flags.resize(quads.size());
  for (size_t q = 0; q < quads.size(); q++) {
    qi[q] = quad_info(quads[q]);
    std::memcpy(&sc.quad_flags[q], &quads[q].p0[3], 4);
  }
  sc.qinfo.upload(qi.data(), qi.size(), b->stream);
  const std::vector<uint32_t> faces = model_faces(v.models, v.models_len, qi, sc.quad_flags);
  sc.mfaces.upload(faces.data(), faces.size(), b->stream);
  sc.atlas.upload(v.atlas, v.atlas_len, b->stream);
  sc.sky.upload(v.sky, v.sky ? v.sky_len : 0, b->stream);
  float lut[256];
  gamma_lut(lut);
  sc.lut.upload(lut, 256, b->stream);

  const lwb::Input in{{v.dims[0], v.dims[1], v.dims[2]},
                      vox.ptr,
                      v.brick_occludes ? occ.ptr : nullptr,
                      flags.ptr,
                      sc.models.ptr,
                      (uint32_t)v.models_len,
                      v.entity_quads_len ? sc.quads.ptr + v.quads_len : nullptr,
                      (uint32_t)v.entity_quads_len,
                      (uint32_t)v.quads_len};
  lwb::Output o;
  try {
    lwb::build(b->stream, in, o);  // waits for the stream: the host vectors above may go
  } catch (const lwb::Error &e) {
    if (e.invalid) throw InvalidScene{e.what};
    throw HipError{e.code, e.what};
  }
  sc.grid.adopt(o.grid, o.grid_len);
  sc.emitters.adopt(o.emitters, o.emitters_len);
  sc.egrid.adopt(o.egrid, o.egrid_len);
  sc.occupancy.adopt(o.occupancy, o.occupancy_len);
  if (!buffers_fit_32bit(o.grid_len, v.models_len, quads.size(), v.atlas_len, o.emitters_len, o.egrid_len,
                         v.sky ? v.sky_len : 0))
    throw Unsupported{"a scene buffer of 4 GiB or more"};
  build_octant_boxes(b, sc, v.dims, o.offsets[0], o.offsets[2]);
  build_cell_ext(b, sc, o.offsets[1], o.bricks);
  HIP_TRY(hipStreamSynchronize(b->stream));
  uint64_t max_faces = 0;
  for (size_t m = 0; m < v.models_len; m++)
    if (v.model_flags[m] & LW_MODEL_EMITS) max_faces = std::max<uint64_t>(max_faces, v.models[m].quad_count);
  sc.max_emitter_rays = (uint32_t)std::min<uint64_t>(std::max<uint64_t>(max_faces * o.max_cell_list, 1), 1u << 20);

  std::memset(&info, 0, sizeof(info));
  for (int a = 0; a < 3; a++) info.dims[a] = v.dims[a];
  std::memcpy(info.offsets, o.offsets, sizeof(info.offsets));
  info.atlas_size[0] = v.atlas_size[0];
  info.atlas_size[1] = v.atlas_size[1];
  info.quads_len = (uint32_t)quads.size();
  info.models_len = (uint32_t)v.models_len;
  info.emitter_count = (uint32_t)o.emitters_count;
  info.see_through[0] = (uint32_t)(v.quads_len + v.entity_quads_len - v.see_through_count);
  info.see_through[1] = v.see_through_count;
  info.has_sun_sprite = v.has_sun_sprite;
  std::memcpy(info.sun_sprite, v.sun_sprite, sizeof(info.sun_sprite));
  sc.is_built = true;
  sc.built = info;
  sc.atlas_len = v.atlas_len;
  sc.sky_len = v.sky ? v.sky_len : 0;
  sc.key_set = false;
  if (stats) {
    std::memset(stats, 0, sizeof(*stats));
    stats->bricks = o.bricks;
    stats->uniform_bricks = o.uniform_bricks;
    stats->buried_bricks = o.buried_bricks;
    stats->instances = v.models_len - 1;
    stats->overlay_bricks = o.overlay_bricks;
    stats->entity_quads = v.entity_quads_len;
    stats->emitters = o.emitters_count;
    stats->grid_len = o.grid_len;
    stats->quads_len = quads.size();
    stats->emitters_len = o.emitters_len;
    stats->emitter_grid_len = o.egrid_len;
    stats->kernel_ms = o.kernel_ms;
    stats->wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }
}

// Launches with the register budget b->waves selects (LW_WAVES: the minimum waves per SIMD
// the compiler is held to; 1 leaves it free).
#define LW_WAVES_DISPATCH(waves, KERNEL_EXPR)  \
  do {                                         \
    switch (waves) {                           \
      case 1: { constexpr int W = 1; KERNEL_EXPR; } break; \
      case 3: { constexpr int W = 3; KERNEL_EXPR; } break; \
      default: { constexpr int W = 4; KERNEL_EXPR; } break; \
    }                                          \
  } while (0)

// The direct kernels' budgets: also 5, 6 and 8 waves per SIMD (96, 80 and 64 VGPRs).
#define LW_DIRECT_DISPATCH(waves, KERNEL_EXPR) \
  do {                                         \
    switch (waves) {                           \
      case 3: { constexpr int W = 3; KERNEL_EXPR; } break; \
      case 4: { constexpr int W = 4; KERNEL_EXPR; } break; \
      case 5: { constexpr int W = 5; KERNEL_EXPR; } break; \
      case 6: { constexpr int W = 6; KERNEL_EXPR; } break; \
      case 8: { constexpr int W = 8; KERNEL_EXPR; } break; \
      default: { constexpr int W = 1; KERNEL_EXPR; } break; \
    }                                          \
  } while (0)

// k_render_tiles<W, SH>: 4, 5 or 6 waves per SIMD (set_tiles_waves keeps the budget in that
// range) and the tile shape SH (0, 1 or 2).
#define LW_TILES_DISPATCH(waves, shape, KERNEL_EXPR)                           \
  do {                                                                        \
    switch ((waves) * 3 + (shape)) {                                          \
      case 12: { constexpr int W = 4, SH = 0; KERNEL_EXPR; } break;           \
      case 13: { constexpr int W = 4, SH = 1; KERNEL_EXPR; } break;           \
      case 14: { constexpr int W = 4, SH = 2; KERNEL_EXPR; } break;           \
      case 15: { constexpr int W = 5, SH = 0; KERNEL_EXPR; } break;           \
      case 16: { constexpr int W = 5, SH = 1; KERNEL_EXPR; } break;           \
      case 17: { constexpr int W = 5, SH = 2; KERNEL_EXPR; } break;           \
      case 18: { constexpr int W = 6, SH = 0; KERNEL_EXPR; } break;           \
      case 19: { constexpr int W = 6, SH = 1; KERNEL_EXPR; } break;           \
      default: { constexpr int W = 6, SH = 2; KERNEL_EXPR; } break;           \
    }                                                                         \
  } while (0)

// A launch layout L (launch_lane) from its runtime value.
#define LW_LAYOUT_DISPATCH(lay, KERNEL_EXPR)                                  \
  do {                                                                        \
    if ((lay) == 2u) {                                                        \
      constexpr int L = 2;                                                    \
      KERNEL_EXPR;                                                            \
    } else if ((lay) == 1u) {                                                 \
      constexpr int L = 1;                                                    \
      KERNEL_EXPR;                                                            \
    } else {                                                                  \
      constexpr int L = 0;                                                    \
      KERNEL_EXPR;                                                            \
    }                                                                         \
  } while (0)

// The differentiable kernels: 4 waves per SIMD, or free (LW_WAVES=1).
#define LW_DIFF_DISPATCH(waves, KERNEL_EXPR)   \
  do {                                         \
    if ((waves) == 1) {                        \
      constexpr int W = 1;                     \
      KERNEL_EXPR;                             \
    } else {                                   \
      constexpr int W = 4;                     \
      KERNEL_EXPR;                             \
    }                                          \
  } while (0)

enum { MODE_PERSISTENT = 0, MODE_DIRECT = 1, MODE_WAVEFRONT = 2 };

// The wavefront loop (k_wf_*) over `work`: iterations until no slot has a ray. Bounded: a
// path has at most MAX_EVENTS events, each with one extension ray, at most 8 shadow
// segments and the scene's emitter rays, and each slot runs at most ceil(total / n) + 1
// paths; past that bound the loop stops with an error instead of spinning.
template <class Work>
void run_wavefront(lw_hip_backend *b, const lw_hip_scene &sc, const lw_params &p, const Work &work,
                   uint32_t total) {
  const uint32_t n = std::max<uint32_t>(
      kBlock, std::min<uint32_t>(b->wf_slots, (uint32_t)(((uint64_t)total + kBlock - 1) / kBlock * kBlock)));
  b->wf_st.ensure(n);
  b->wf_o.ensure(n);
  b->wf_d.ensure(n);
  b->wf_hit.ensure(n);
  b->wf_queue.ensure(n);
  b->wf_ctr.ensure(4);
  const WfSlots w{b->wf_st.ptr, b->wf_o.ptr, b->wf_d.ptr, b->wf_hit.ptr, b->wf_queue.ptr,
                  b->wf_ctr.ptr, n};
  const DevScene ds = sc.dev();
  const dim3 grid(n / kBlock), block(kBlock);
  k_wf_init<<<grid, block, 0, b->stream>>>(w);
  HIP_TRY(hipGetLastError());
  const uint64_t per_path = (uint64_t)MAX_EVENTS * (1 + 8 + sc.max_emitter_rays);
  const uint64_t limit = ((uint64_t)total / n + 2) * per_path + 64;
  uint64_t it = 0;
  for (;; it++) {
    if (it > limit) throw InvalidScene{"wavefront loop exceeded its iteration bound"};
    LW_WAVES_DISPATCH(b->waves, (k_wf_logic<Work, W><<<grid, block, 0, b->stream>>>(p, ds, work, w)));
    k_wf_flip<<<dim3(1), dim3(64), 0, b->stream>>>(w.ctr);
    if (b->trace_waves >= 8)
      k_wf_trace<8><<<grid, block, 0, b->stream>>>(p, ds, w);
    else
      k_wf_trace<1><<<grid, block, 0, b->stream>>>(p, ds, w);
    HIP_TRY(hipGetLastError());
    if ((it & 7) == 7) {
      HIP_TRY(hipMemcpyAsync(b->pinned, w.ctr + 1, sizeof(uint32_t), hipMemcpyDeviceToHost, b->stream));
      HIP_TRY(hipStreamSynchronize(b->stream));
      if (*b->pinned == 0) break;
    }
  }
  b->last_iterations = it + 1;
}

constexpr double kTargetLaunchMs = 150.0;  // render/cuda.rs:19
// Frames with at least this many samples per pixel render in one persistent k_render_tiles
// launch (a wave's sample slots, 16 or more per pixel, are then busy in every round).
constexpr uint32_t kTilesMinSpp = 16;
// ... when it has an item for every resident wave and at least kTilesPerWave rounds (a round
// being one wave's 64 samples of its tile) for each: then the last rounds to start are a small
// part of the launch (measured with 2 x 2-pixel tiles: 256 x 144 at 64 spp, 7.2 rounds a wave,
// 712 Msamples/s persistent against 664 in launches; 320 x 180 at 16 spp, 2.8 rounds a wave,
// 940 against 986).
constexpr uint32_t kTilesPerWave = 4;

// The tile kernel's budget and the resident workgroups it allows (its launch's grid).
void set_tiles_waves(lw_hip_backend *b, int waves) {
  b->tiles_waves = std::min(std::max(waves, 4), 6);
  hipDeviceProp_t prop;
  HIP_TRY(hipGetDeviceProperties(&prop, b->device));
  int per_cu = 0;
  LW_TILES_DISPATCH(b->tiles_waves, 2, {
    HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &per_cu, reinterpret_cast<const void *>(&k_render_tiles<W, SH>), kBlock, 0));
  });
  b->blocks_tiles = (uint32_t)std::max(1, prop.multiProcessorCount) * (uint32_t)std::max(1, per_cu);
}

// A view's tile shape in k_render_tiles: the fewest pixels per wave (one pixel's 64 samples a
// round, else two pixels' 32) whose rounds leave at most an eighth of the wave's slots idle,
// else 2 x 2 pixels' 16; and a larger tile where the view would have 2^24 tiles or more (an
// item holds the tile in 24 bits). 256 spp: one pixel; 84: two; 100 or 16: 2 x 2.
uint32_t tile_shape(uint32_t w, uint32_t h, uint32_t spp) {
  uint32_t sh = 2;
  for (uint32_t c = 0; c < 2; c++) {
    const uint64_t ns = 64u >> c, slots = (spp + ns - 1) / ns * ns;
    if ((slots - spp) * 8 <= slots) {
      sh = c;
      break;
    }
  }
  while (sh < 2 && tile_count(w, h, sh) >= (1u << 24)) sh++;
  return sh;
}

// The tile shape of a batch's one k_render_tiles launch: every view's shape when they agree,
// else the shape of the fewest wave slots per sample over all views, each shape's slots
// weighted by its measured cost per slot (1, 1.015, 1.035 for one, two and four pixels per
// wave); and then a larger tile while a view would have 2^24 tiles or more.
uint32_t batch_tile_shape(const BatchDev &bd) {
  const uint32_t nv = bd.nviews;
  uint32_t common = tile_shape(bd.cam[0].width, bd.cam[0].height, bd.cam[0].spp);
  for (uint32_t v = 1; v < nv && common < 3; v++)
    if (tile_shape(bd.cam[v].width, bd.cam[v].height, bd.cam[v].spp) != common) common = 3;
  uint32_t sh = common;
  if (common == 3) {
    const double weight[3] = {1.0, 1.015, 1.035};
    double best = 0.0;
    for (uint32_t c = 0; c < 3; c++) {
      const uint64_t ns = 64u >> c;
      double slots = 0.0;
      for (uint32_t v = 0; v < nv; v++) {
        const ViewCam &cam = bd.cam[v];
        slots += (double)cam.width * cam.height * (double)((cam.spp + ns - 1) / ns * ns);
      }
      if (c == 0 || slots * weight[c] < best) {
        best = slots * weight[c];
        sh = c;
      }
    }
  }
  for (bool big = true; big && sh < 2;) {
    big = false;
    for (uint32_t v = 0; v < nv; v++)
      if (tile_count(bd.cam[v].width, bd.cam[v].height, sh) >= (1u << 24)) big = true;
    if (big) sh++;
  }
  return sh;
}

// Whether a frame renders in one persistent tile launch (render_frame's rule).
bool uses_tiles(const lw_hip_backend *b, uint32_t w, uint32_t h, uint32_t spp) {
  const uint32_t sh = tile_shape(w, h, spp);
  const uint64_t tile_items = tile_count(w, h, sh);
  const uint64_t waves = (uint64_t)b->blocks_tiles * (kBlock / 64);
  const uint64_t rounds = tile_items * ((spp + (64u >> sh) - 1) / (64u >> sh));
  // (a frame of 2^24 tiles or more, which an item cannot name, takes the launch ramp)
  return b->mode == MODE_DIRECT && b->per_lane == 1 && !b->frame_launches && spp >= kTilesMinSpp &&
         tile_items >= waves && rounds >= (uint64_t)kTilesPerWave * waves && tile_items < (1u << 24);
}

// A frame in one persistent launch (k_render_tiles over its tiles, all samples of a tile in one
// wave, summed in sample order in f64): the image equals the launch ramp's bit for bit. With `progress`, the host reports (samples done, spp) while the launch runs, from the item
// index the kernel last stored to host memory, and (spp, spp) at the end.
void render_frame_tiles(lw_hip_backend *b, const lw_hip_scene &sc, const lw_params &p, uint32_t spp,
                        lw_progress_fn progress, void *user, float *out_rgb,
                        std::chrono::steady_clock::time_point t_wall) {
  const uint32_t w = p.frame[0], h = p.frame[1];
  const uint64_t npix = (uint64_t)w * h;
  if (npix > (1ull << 26)) throw InvalidScene{"frame too large"};
  const uint32_t sh = tile_shape(w, h, spp);
  const uint64_t n_tiles = tile_count(w, h, sh);
  if (n_tiles >= (1u << 24)) throw InvalidScene{"frame too large"};
  const uint32_t n_items = (uint32_t)n_tiles;
  BatchPlan plan;
  BatchDev &bd = plan.bd();
  std::memset(&bd, 0, sizeof(bd));
  ViewCam &c = bd.cam[0];
  std::memcpy(c.cam_pos, p.cam_pos, sizeof(c.cam_pos));
  std::memcpy(c.cam_fwd, p.cam_fwd, sizeof(c.cam_fwd));
  std::memcpy(c.cam_right, p.cam_right, sizeof(c.cam_right));
  std::memcpy(c.cam_up, p.cam_up, sizeof(c.cam_up));
  c.width = w;
  c.height = h;
  c.seed = p.misc_u[0];
  c.spp = spp;
  bd.nviews = 1;
  bd.pix_prefix[1] = (uint32_t)npix;
  bd.npix_total = (uint32_t)npix;
  b->batch_dev.ensure(1);
  b->out.ensure(npix * 3);
  HIP_TRY(hipMemcpyAsync(b->batch_dev.ptr, &bd, sizeof(BatchDev), hipMemcpyHostToDevice, b->stream));
  k_fill_u32<<<dim3(1), dim3(64), 0, b->stream>>>(b->counter.ptr, 0u);
  HIP_TRY(hipGetLastError());
  if (progress) __atomic_store_n(b->progress_host, 0u, __ATOMIC_RELAXED);
  const DevScene ds = sc.dev();
  const uint32_t grid = std::max<uint32_t>(1, std::min<uint32_t>(b->blocks_tiles, (n_items + 3) / 4));
  HIP_TRY(hipEventRecord(b->ev[0], b->stream));
  LW_TILES_DISPATCH(b->tiles_waves, sh, (k_render_tiles<W, SH><<<dim3(grid), dim3(kBlock), 0, b->stream>>>(
                                   p, ds, b->batch_dev.ptr, nullptr, n_items, b->counter.ptr, b->out.ptr,
                                   progress ? b->progress_dev : nullptr)));
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipEventRecord(b->ev[1], b->stream));
  if (progress) {
    uint32_t reported = 0;
    for (;;) {
      const hipError_t q = hipEventQuery(b->ev[1]);
      if (q == hipSuccess) break;
      if (q != hipErrorNotReady) throw HipError{q, "hipEventQuery"};
      std::this_thread::sleep_for(std::chrono::milliseconds(20));
      // every item before the one last stored has been taken; count those samples as done
      const uint64_t taken = __atomic_load_n(b->progress_host, __ATOMIC_RELAXED);
      const uint32_t done = (uint32_t)std::min<uint64_t>(spp - 1, taken * spp / n_items);
      if (done > reported) {
        reported = done;
        progress(done, spp, user);
      }
    }
  }
  if (out_rgb)  // without it the image stays in b->out for the caller's next kernel
    HIP_TRY(hipMemcpyAsync(out_rgb, b->out.ptr, npix * 3 * sizeof(float), hipMemcpyDeviceToHost, b->stream));
  HIP_TRY(hipStreamSynchronize(b->stream));
  if (progress) progress(spp, spp, user);
  float ms = 0.0f;
  HIP_TRY(hipEventElapsedTime(&ms, b->ev[0], b->ev[1]));
  lw_timing tm{};
  tm.trace_ms = ms;
  tm.launches = 1;
  tm.samples = npix * spp;
  tm.wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_wall).count();
  b->timing = tm;
}
constexpr uint64_t kFillBlocks = 2048;     // 8 workgroups of 256 lanes per CU
constexpr uint32_t kMaxBatch = 64;         // render/cuda.rs:165

// Device memory a frame renders into when the caller provides it (FrameScratch) instead of
// the backend's grow-only buffers: the f64 accumulator and `rad_floats` floats of staging,
// which also take the f32 means at the end. Such a frame runs the direct form.
struct FrameScratch {
  double *acc;
  float *rad;
  uint64_t rad_floats;
};

// The scratch array's split for a frame of npix pixels: the accumulator at its start, the
// staging after it, 256-byte aligned. False when the frame does not fit one sample per launch.
bool split_scratch(void *base, uint64_t npix, FrameScratch &fs) {
  const uint64_t acc_bytes = (npix * 3 * sizeof(double) + 255) / 256 * 256;
  if (acc_bytes + npix * 3 * sizeof(float) > kScratchBytes) return false;
  fs.acc = static_cast<double *>(base);
  fs.rad = reinterpret_cast<float *>(static_cast<unsigned char *>(base) + acc_bytes);
  fs.rad_floats = (kScratchBytes - acc_bytes) / sizeof(float);
  return true;
}

// With `diff`, the samples are render_sample_diff's under its parameters (direct form).
// With `out_f64`, it also receives the per-pixel means in f64 (sum / spp before the f32
// rounding of out_rgb).
// With `one_launch`, all spp samples go in one launch (up to kMaxBatch), for timing probes.
// With `fixed`, the frame allocates nothing: it renders into that memory, in the direct form.
void render_frame(lw_hip_backend *b, const lw_hip_scene &sc, const lw_params &params_in,
                  uint32_t spp_in, lw_progress_fn progress, void *user, float *out_rgb,
                  const DiffDev *diff = nullptr, double *out_f64 = nullptr,
                  bool one_launch = false, const FrameScratch *fixed = nullptr) {
  auto t_wall = std::chrono::steady_clock::now();
  check_layout(sc, params_in);
  lw_params p = params_in;
  p.frame[2] = 0;
  const uint32_t w = p.frame[0], h = p.frame[1];
  if (w == 0 || h == 0) throw InvalidScene{"empty frame"};
  const uint64_t npix = (uint64_t)w * h;
  if (npix > (1ull << 31) / 3) throw InvalidScene{"frame too large"};
  const uint32_t spp = std::max(spp_in, 1u);
  HIP_TRY(hipSetDevice(b->device));
  const DevScene ds = sc.dev();

  // The persistent form needs an item for each resident wave and kTilesPerWave rounds for each,
  // so that the last rounds to start are a small part of the launch; smaller frames fill the
  // device through the launch ramp's many samples per launch.
  if (!diff && !out_f64 && !one_launch && !fixed && uses_tiles(b, w, h, spp)) {
    render_frame_tiles(b, sc, p, spp, progress, user, out_rgb, t_wall);
    return;
  }

  // Staging for up to kMaxBatch samples per launch, capped at 2 GiB (or at what the caller's
  // memory holds).
  uint32_t max_batch = kMaxBatch;
  while (max_batch > 1 && npix * 3 * sizeof(float) * max_batch > (2ull << 30)) max_batch /= 2;
  float *rad, *out;
  double *acc;
  int mode = b->mode;
  if (fixed) {
    while (max_batch > 1 && npix * 3 * max_batch > fixed->rad_floats) max_batch /= 2;
    if (npix * 3 > fixed->rad_floats) throw InvalidScene{"frame too large for the scratch memory"};
    rad = fixed->rad;
    acc = fixed->acc;
    out = fixed->rad;  // the staging is free once the last launch is summed
    mode = MODE_DIRECT;
  } else {
    b->rad.ensure(npix * 3 * max_batch);
    b->acc.ensure(npix * 3);
    b->out.ensure(npix * 3);
    rad = b->rad.ptr;
    acc = b->acc.ptr;
    out = b->out.ptr;
  }
  {
    const size_t n = npix * 3;
    k_zero<<<dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, b->stream>>>(acc, n);
    HIP_TRY(hipGetLastError());
  }

  const uint32_t tiles_x = (w + 15) / 16;
  const uint32_t tiles = tiles_x * ((h + 15) / 16);
  if ((uint64_t)tiles * 256u * max_batch > 0xFFFFFFFFull) throw InvalidScene{"frame too large"};
  lw_timing tm{};
  // render/cuda.rs:131-175 starts at one sample per pixel and doubles while launches stay
  // short. A frame too small to fill the GPU with one sample per pixel starts at the first
  // power of two that gives it kFillBlocks workgroups (same samples, same summation order).
  uint32_t done = 0, batch = 1;
  while (batch < max_batch && (uint64_t)tiles * batch < kFillBlocks) batch *= 2;
  if (one_launch) batch = max_batch;
  while (done < spp) {
    const uint32_t nb = std::min(batch, std::min(spp - done, max_batch));
    const auto start = std::chrono::steady_clock::now();
    HIP_TRY(hipEventRecord(b->ev[0], b->stream));
    if (diff) {
      const uint32_t lay = launch_layout(nb);
      const dim3 grid = launch_grid(lay, w, h, nb);
      LW_LAYOUT_DISPATCH(lay, LW_DIFF_DISPATCH(b->waves, (k_render_diff<W, L><<<grid, dim3(kBlock), 0, b->stream>>>(
                                                              p, ds, *diff, done, rad))));
    } else if (mode == MODE_DIRECT && b->per_lane > 1) {
      const uint32_t per = std::min(b->per_lane, nb);
      LW_DIRECT_DISPATCH(b->waves, (k_render_multi<W><<<dim3(tiles, (nb + per - 1) / per),
                                                         dim3(kBlock), 0, b->stream>>>(
                                       p, ds, done, nb, per, rad)));
    } else if (mode == MODE_DIRECT && nb % 16u == 0 && !b->xcd_remap) {
      // a wave per pixel's 64 samples, or 2 x 2 pixels' 16
      const uint32_t lay = launch_layout(nb);
      const dim3 grid = launch_grid(lay, w, h, nb);
      if (lay == 2u) {
        LW_DIRECT_DISPATCH(b->waves, (k_render_slots<W, 2><<<grid, dim3(kBlock), 0, b->stream>>>(
                                         p, ds, done, rad)));
      } else {
        LW_DIRECT_DISPATCH(b->waves, (k_render_slots<W, 1><<<grid, dim3(kBlock), 0, b->stream>>>(
                                         p, ds, done, rad)));
      }
    } else if (mode == MODE_DIRECT) {
      LW_DIRECT_DISPATCH(b->waves, (k_render<W><<<dim3(tiles, nb), dim3(kBlock), 0, b->stream>>>(
                                       p, ds, done, b->xcd_remap, rad)));
    } else if (mode == MODE_WAVEFRONT) {
      const FrameWork fw{w, h, tiles_x, tiles, done, tiles * 256u * nb, rad};
      run_wavefront(b, sc, p, fw, fw.total);
    } else {
      const FrameWork fw{w, h, tiles_x, tiles, done, tiles * 256u * nb, rad};
      const uint32_t grid = std::min<uint32_t>(b->blocks_frame, (fw.total + kBlock - 1) / kBlock);
      k_fill_u32<<<dim3(1), dim3(64), 0, b->stream>>>(b->counter.ptr, 0u);
      LW_WAVES_DISPATCH(b->waves, (k_paths<FrameWork, W><<<dim3(grid), dim3(kBlock), 0, b->stream>>>(
                                      p, ds, fw, b->counter.ptr)));
    }
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(b->ev[1], b->stream));
    k_accumulate<<<dim3((unsigned)((npix + kBlock - 1) / kBlock)), dim3(kBlock), 0, b->stream>>>(
        (uint32_t)npix, nb, rad, acc);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(b->ev[2], b->stream));
    HIP_TRY(hipEventSynchronize(b->ev[2]));
    float ms_r = 0.0f, ms_a = 0.0f;
    HIP_TRY(hipEventElapsedTime(&ms_r, b->ev[0], b->ev[1]));
    HIP_TRY(hipEventElapsedTime(&ms_a, b->ev[1], b->ev[2]));
    tm.trace_ms += ms_r;
    tm.accumulate_ms += ms_a;
    tm.launches += 1;
    done += nb;
    if (progress) progress(done, spp, user);
    const double took =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - start).count();
    if (took < kTargetLaunchMs / 2)
      batch = std::min(batch * 2, kMaxBatch);
    else if (took > kTargetLaunchMs * 2 && batch > 1)
      batch /= 2;
  }
  {
    const size_t n = npix * 3;
    k_resolve<<<dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, b->stream>>>(
        (uint32_t)npix, spp, acc, out);
    HIP_TRY(hipGetLastError());
  }
  if (out_rgb)  // without it the image stays in b->out for the caller's next kernel
    HIP_TRY(hipMemcpyAsync(out_rgb, out, npix * 3 * sizeof(float), hipMemcpyDeviceToHost,
                           b->stream));
  if (out_f64)
    HIP_TRY(hipMemcpyAsync(out_f64, acc, npix * 3 * sizeof(double), hipMemcpyDeviceToHost,
                           b->stream));
  HIP_TRY(hipStreamSynchronize(b->stream));
  if (out_f64)
    for (size_t i = 0; i < npix * 3; i++) out_f64[i] /= (double)spp;
  tm.samples = npix * spp;
  tm.wall_ms =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_wall).count();
  b->timing = tm;
}

// The views of a batch share every Params field but the camera, the frame size and the
// seed (the kernels read those per view from BatchDev, the rest from the first view).
void check_batch_view(const lw_params &a, const lw_params &v) {
  lw_params x = v;
  std::memcpy(x.cam_pos, a.cam_pos, sizeof(x.cam_pos));
  std::memcpy(x.cam_fwd, a.cam_fwd, sizeof(x.cam_fwd));
  std::memcpy(x.cam_right, a.cam_right, sizeof(x.cam_right));
  std::memcpy(x.cam_up, a.cam_up, sizeof(x.cam_up));
  x.frame[0] = a.frame[0];
  x.frame[1] = a.frame[1];
  x.frame[2] = a.frame[2];
  x.misc_u[0] = a.misc_u[0];
  if (std::memcmp(&x, &a, sizeof(x)) != 0)
    throw InvalidScene{"batch views differ in more than camera, frame size and seed"};
}

BatchPlan plan_batch(lw_hip_backend *b, const lw_hip_scene &sc, const lw_view *views, size_t nv) {
  if (nv == 0 || nv > kMaxViews) throw InvalidScene{"a batch holds 1 to 256 views"};
  const lw_params &p0 = views[0].params;
  check_layout(sc, p0);
  BatchPlan plan;
  BatchDev &bd = plan.bd();
  std::memset(&bd, 0, sizeof(bd));
  bd.nviews = (uint32_t)nv;
  uint64_t npix_total = 0, samples = 0, max_spp = 0, staging_all = 0;
  for (size_t v = 0; v < nv; v++) {
    const lw_params &p = views[v].params;
    check_batch_view(p0, p);
    const uint32_t w = p.frame[0], h = p.frame[1], spp = std::max(views[v].spp, 1u);
    if (w == 0 || h == 0 || (uint64_t)w * h > (1u << 26)) throw InvalidScene{"bad view size"};
    ViewCam &c = bd.cam[v];
    std::memcpy(c.cam_pos, p.cam_pos, sizeof(c.cam_pos));
    std::memcpy(c.cam_fwd, p.cam_fwd, sizeof(c.cam_fwd));
    std::memcpy(c.cam_right, p.cam_right, sizeof(c.cam_right));
    std::memcpy(c.cam_up, p.cam_up, sizeof(c.cam_up));
    c.width = w;
    c.height = h;
    c.seed = p.misc_u[0];
    c.spp = spp;
    bd.tiles_x[v] = (w + 15) / 16;
    bd.tiles[v] = bd.tiles_x[v] * ((h + 15) / 16);
    bd.pix_prefix[v] = (uint32_t)npix_total;
    npix_total += (uint64_t)w * h;
    samples += (uint64_t)w * h * spp;
    staging_all += (uint64_t)w * h * spp * 3 * sizeof(float);
    max_spp = std::max<uint64_t>(max_spp, spp);
  }
  if (npix_total * 3 > 0xFFFFFFFFull) throw InvalidScene{"batch too large"};
  bd.pix_prefix[nv] = (uint32_t)npix_total;
  bd.npix_total = (uint32_t)npix_total;
  // Passes: as few as fit the staging cap and keep a pass's items below 2^31.
  uint64_t passes = std::max<uint64_t>(1, (staging_all + b->batch_staging - 1) / b->batch_staging);
  if (b->batch_passes) passes = std::max<uint64_t>(passes, b->batch_passes);
  passes = std::min<uint64_t>(passes, max_spp);
  uint64_t staging = 0, items = 0;
  for (;;) {
    staging = 0;
    items = 0;
    for (size_t v = 0; v < nv; v++) {
      // a multiple of 16 where the view has that many, for 2 x 2-pixel x 16-sample waves
      const uint64_t spp_v = bd.cam[v].spp, m0 = (spp_v + passes - 1) / passes;
      const uint32_t m = (uint32_t)(m0 >= 16 ? std::min<uint64_t>((m0 + 15) / 16 * 16, spp_v) : m0);
      bd.per_pass[v] = m;
      bd.rad_base[v] = staging / sizeof(float);
      staging += (uint64_t)bd.cam[v].width * bd.cam[v].height * m * 3 * sizeof(float);
      items += (uint64_t)bd.tiles[v] * 256 * m;
    }
    if ((staging <= b->batch_staging && items < (1ull << 31)) || passes >= max_spp) break;
    passes++;
  }
  uint32_t npass = 0;
  for (size_t v = 0; v < nv; v++)
    npass = std::max<uint32_t>(npass, (bd.cam[v].spp + bd.per_pass[v] - 1) / bd.per_pass[v]);
  if (items >= (1ull << 31)) throw InvalidScene{"batch too large for its passes"};
  plan.npass = npass;
  plan.items = items;
  plan.staging = staging;
  plan.npix_total = npix_total;
  plan.samples = samples;
  return plan;
}

// Whether every pass of a planned batch renders a multiple of 64 samples of each view (its
// samples per pass and its spp both multiples of 64), so that its passes can run a wave per
// pixel (batch_item<2>).
bool batch_by64(const BatchDev &bd) {
  for (uint32_t v = 0; v < bd.nviews; v++)
    if (bd.per_pass[v] % 64u != 0 || bd.cam[v].spp % 64u != 0) return false;
  return bd.nviews > 0;
}

// With `diff`, the samples are render_sample_diff's (direct form). Without `outs`, the views'
// f64 sums stay in b->acc (laid out by pix_prefix) for the caller.
void render_batch(lw_hip_backend *b, const lw_hip_scene &sc, const lw_view *views, size_t nv,
                  lw_progress_fn progress, void *user, float *const *outs,
                  const DiffDev *diff = nullptr) {
  auto t_wall = std::chrono::steady_clock::now();
  if (outs)
    for (size_t v = 0; v < nv; v++)
      if (!outs[v]) throw InvalidScene{"null output for a view"};
  b->plan = plan_batch(b, sc, views, nv);
  const BatchPlan &plan = b->plan;
  const BatchDev &bd = plan.bd();
  const lw_params &p0 = views[0].params;
  const uint32_t npass = plan.npass;
  const uint64_t items = plan.items, staging = plan.staging, npix_total = plan.npix_total,
                 samples = plan.samples;
  HIP_TRY(hipSetDevice(b->device));
  const DevScene ds = sc.dev();
  const bool tiles_form = !diff && outs && b->mode == MODE_DIRECT && !b->batch_passes_form;
  b->batch_dev.ensure(1);
  if (!tiles_form) {
    b->rad.ensure(staging / sizeof(float) + 1);
    b->acc.ensure(npix_total * 3);
  }
  b->out.ensure(npix_total * 3);
  HIP_TRY(hipMemcpyAsync(b->batch_dev.ptr, &bd, sizeof(BatchDev), hipMemcpyHostToDevice, b->stream));
  if (!tiles_form) {
    k_zero<<<dim3((unsigned)((npix_total * 3 + kBlock - 1) / kBlock)), dim3(kBlock), 0, b->stream>>>(
        b->acc.ptr, npix_total * 3);
    HIP_TRY(hipGetLastError());
  }
  const BatchWork work{b->batch_dev.ptr, b->rad.ptr};
  const dim3 acc_grid((unsigned)((npix_total + kBlock - 1) / kBlock));
  // Pass 0 has the most items (a view's samples per pass only shrink at its last pass).
  const uint32_t direct_blocks = (uint32_t)std::max<uint64_t>(1, items / kBlock);
  auto pass_ops = [&](hipStream_t st) {
    k_batch_plan<<<dim3(1), dim3(kMaxViews), 0, st>>>(b->batch_dev.ptr);
    k_fill_u32<<<dim3(1), dim3(64), 0, st>>>(b->counter.ptr, 0u);
    if (diff) {
      if (batch_by64(bd)) {
        LW_DIFF_DISPATCH(b->waves, (k_render_diff_batch<W, 2><<<dim3(direct_blocks), dim3(kBlock), 0, st>>>(
                                       p0, ds, *diff, b->batch_dev.ptr, b->rad.ptr)));
      } else {
        LW_DIFF_DISPATCH(b->waves, (k_render_diff_batch<W, 0><<<dim3(direct_blocks), dim3(kBlock), 0, st>>>(
                                       p0, ds, *diff, b->batch_dev.ptr, b->rad.ptr)));
      }
    } else if (b->mode == MODE_DIRECT) {
      LW_DIRECT_DISPATCH(b->waves, (k_render_batch<W><<<dim3(direct_blocks), dim3(kBlock), 0, st>>>(
                                       p0, ds, b->batch_dev.ptr, b->rad.ptr)));
    } else {
      LW_WAVES_DISPATCH(b->waves, (k_paths<BatchWork, W><<<dim3(b->blocks_frame), dim3(kBlock), 0, st>>>(
                                      p0, ds, work, b->counter.ptr)));
    }
    k_accumulate_batch<<<acc_grid, dim3(kBlock), 0, st>>>(b->batch_dev.ptr, b->rad.ptr, b->acc.ptr);
  };
  lw_timing tm{};
  HIP_TRY(hipEventRecord(b->ev[0], b->stream));
  if (tiles_form) {
    // One persistent launch over (view, tile) items, each view with its own tile shape, the
    // views with the most samples per pixel first so that the last items to start are the
    // shortest.
    std::vector<uint32_t> order(nv);
    for (size_t v = 0; v < nv; v++) order[v] = (uint32_t)v;
    std::stable_sort(order.begin(), order.end(),
                     [&](uint32_t a, uint32_t c) { return bd.cam[a].spp > bd.cam[c].spp; });
    const uint32_t shape = batch_tile_shape(bd);
    std::vector<uint32_t> list;
    for (uint32_t v : order) {
      const uint64_t n = tile_count(bd.cam[v].width, bd.cam[v].height, shape);
      if (n >= (1u << 24)) throw InvalidScene{"a view of the batch has 2^24 tiles or more"};
      for (uint32_t t = 0; t < (uint32_t)n; t++) list.push_back(v << 24 | t);
    }
    b->batch_items.upload(list.data(), list.size(), b->stream);
    k_fill_u32<<<dim3(1), dim3(64), 0, b->stream>>>(b->counter.ptr, 0u);
    const uint32_t grid = std::max<uint32_t>(1, std::min<uint32_t>(b->blocks_tiles, (uint32_t)((list.size() + 3) / 4)));
    LW_TILES_DISPATCH(b->tiles_waves, shape, (k_render_tiles<W, SH><<<dim3(grid), dim3(kBlock), 0, b->stream>>>(
                                     p0, ds, b->batch_dev.ptr, b->batch_items.ptr, (uint32_t)list.size(),
                                     b->counter.ptr, b->out.ptr, nullptr)));
    HIP_TRY(hipGetLastError());
    if (progress) {
      HIP_TRY(hipStreamSynchronize(b->stream));
      progress(1, 1, user);
    }
  } else if (b->mode == MODE_WAVEFRONT && !diff) {
    // The wavefront loop's length depends on the pass, so each pass is launched from here.
    for (uint32_t k = 0; k < npass; k++) {
      k_batch_plan<<<dim3(1), dim3(kMaxViews), 0, b->stream>>>(b->batch_dev.ptr);
      run_wavefront(b, sc, p0, work, (uint32_t)items);
      k_accumulate_batch<<<acc_grid, dim3(kBlock), 0, b->stream>>>(b->batch_dev.ptr, b->rad.ptr,
                                                                    b->acc.ptr);
      HIP_TRY(hipGetLastError());
      if (progress) {
        HIP_TRY(hipStreamSynchronize(b->stream));
        progress(k + 1, npass, user);
      }
    }
  } else if (b->use_graphs) {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    pass_ops(b->stream);
    HIP_TRY(hipStreamEndCapture(b->stream, &graph));
    hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
      (void)hipGraphDestroy(graph);
      throw HipError{e, "hipGraphInstantiate"};
    }
    for (uint32_t k = 0; k < npass; k++) {
      e = hipGraphLaunch(exec, b->stream);
      if (e == hipSuccess && progress) {
        e = hipStreamSynchronize(b->stream);
        if (e == hipSuccess) progress(k + 1, npass, user);
      }
      if (e != hipSuccess) break;
    }
    (void)hipStreamSynchronize(b->stream);
    (void)hipGraphExecDestroy(exec);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) throw HipError{e, "hipGraphLaunch"};
  } else {
    for (uint32_t k = 0; k < npass; k++) {
      pass_ops(b->stream);
      HIP_TRY(hipGetLastError());
      if (progress) {
        HIP_TRY(hipStreamSynchronize(b->stream));
        progress(k + 1, npass, user);
      }
    }
  }
  HIP_TRY(hipEventRecord(b->ev[1], b->stream));
  if (outs) {
    if (!tiles_form) {
      k_resolve_batch<<<dim3((unsigned)((npix_total * 3 + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                        b->stream>>>(b->batch_dev.ptr, b->acc.ptr, b->out.ptr);
      HIP_TRY(hipGetLastError());
    }
    for (size_t v = 0; v < nv; v++) {
      const size_t n3 = (size_t)bd.cam[v].width * bd.cam[v].height * 3;
      HIP_TRY(hipMemcpyAsync(outs[v], b->out.ptr + (size_t)bd.pix_prefix[v] * 3, n3 * sizeof(float),
                             hipMemcpyDeviceToHost, b->stream));
    }
  }
  HIP_TRY(hipStreamSynchronize(b->stream));
  float ms = 0.0f;
  HIP_TRY(hipEventElapsedTime(&ms, b->ev[0], b->ev[1]));
  tm.trace_ms = ms;
  tm.launches = tiles_form ? 1u : npass;
  tm.samples = samples;
  tm.wall_ms =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_wall).count();
  b->timing = tm;
}

// ---- differentiable rendering: argument checks, device copies, the adjoint loop ----

constexpr uint32_t kMaxDiffSlots = 6144;  // f64 gradient slots a workgroup keeps in LDS (48 KiB)

// Checks a lw_diff_desc against the scene and params and uploads its arrays; the returned
// DiffDev points into b's buffers.
DiffDev prepare_diff(lw_hip_backend *b, const lw_hip_scene &sc, const lw_params &p,
                     const lw_diff_desc &dd) {
  auto fail = [](const std::string &m) { throw InvalidScene{m}; };
  const uint64_t slots = 3ull * dd.n_palette + dd.n_intensity + 3ull;
  if (slots > kMaxDiffSlots) fail("too many palette and intensity groups");
  const size_t nq = sc.quad_flags.size();
  if (!dd.quad_palette || !dd.quad_intensity) fail("quad group arrays are null");
  if ((dd.n_palette && !dd.palette) || (dd.n_intensity && !dd.intensity)) fail("factors are null");
  for (size_t q = 0; q < nq; q++) {
    const uint32_t g = dd.quad_palette[q], ig = dd.quad_intensity[q];
    if (g != LW_NO_GROUP && g >= dd.n_palette) fail("palette group out of range");
    if (ig != LW_NO_GROUP && ig >= dd.n_intensity) fail("intensity group out of range");
    if (g != LW_NO_GROUP && (sc.quad_flags[q] & LW_FLAG_ALPHA_MASK) == LW_ALPHA_TRANSLUCENT)
      fail("a translucent quad has a palette group");
  }
  for (uint32_t i = 0; i < 3 * dd.n_palette; i++)
    if (!(std::isfinite(dd.palette[i]) && dd.palette[i] > 0.0f)) fail("palette factors must be finite and > 0");
  for (uint32_t i = 0; i < dd.n_intensity; i++)
    if (!(std::isfinite(dd.intensity[i]) && dd.intensity[i] > 0.0f)) fail("intensity factors must be finite and > 0");
  const uint32_t strategy = p.misc_u[3] > 0 ? p.emitter[0] : LW_EMITTER_SAMPLING_NONE;
  if (strategy != LW_EMITTER_SAMPLING_NONE && strategy != LW_EMITTER_SAMPLING_ONE)
    throw Unsupported{"differentiable rendering samples emitters with NONE or ONE only"};
  uint32_t cache_len = 0;
  if (dd.sky_params) {
    if (!(p.sky_map[0] < 1.0f) || !(p.misc_f[2] >= 1.0f))
      throw Unsupported{"sky parameters need the Preetham sky cache (no sky maps)"};
    if (p.sun_sprite[2] > 0.0f) throw Unsupported{"sky parameters with the sun sprite drawn"};
    const uint64_t n = (uint64_t)p.misc_f[2] + 1;
    if (n * n * 3 > sc.sky.n) fail("sky cache shorter than the params' resolution");
    cache_len = (uint32_t)(n * n * 3);
    if (!dd.d_sky_cache) fail("d_sky_cache is null");
  }
  HIP_TRY(hipSetDevice(b->device));
  b->diff_pal.upload(dd.palette, 3 * (size_t)dd.n_palette, b->stream);
  b->diff_inten.upload(dd.intensity, dd.n_intensity, b->stream);
  b->diff_qg.upload(dd.quad_palette, nq, b->stream);
  b->diff_qig.upload(dd.quad_intensity, nq, b->stream);
  b->diff_dsky.upload(dd.d_sky_cache, dd.sky_params ? 3 * (size_t)cache_len : 0, b->stream);
  HIP_TRY(hipStreamSynchronize(b->stream));  // the caller's arrays may go after the call
  DiffDev D{};
  D.pal = b->diff_pal.ptr;
  D.inten = b->diff_inten.ptr;
  D.qgroup = b->diff_qg.ptr;
  D.qigroup = b->diff_qig.ptr;
  D.dsky = b->diff_dsky.ptr;
  for (int k = 0; k < 3; k++) {
    for (int c = 0; c < 3; c++) {
      D.dsun[k][c] = dd.sky_params ? dd.d_sun_dir[k][c] : 0.0f;
      D.dsun_color[k][c] = dd.sky_params ? dd.d_sun_color[k][c] : 0.0f;
    }
    D.dsky_light[k] = dd.sky_params ? dd.d_sky_light[k] : 0.0f;
  }
  D.n_pal = dd.n_palette;
  D.n_int = dd.n_intensity;
  D.rr_off = dd.rr_off ? 1u : 0u;
  D.sky_params = dd.sky_params ? 1u : 0u;
  D.cache_len = cache_len;
  return D;
}

void adjoint_frame(lw_hip_backend *b, const lw_hip_scene &sc, const lw_params &params_in,
                   uint32_t spp_in, const lw_diff_desc &dd, const float *dldi, double *grads) {
  auto t_wall = std::chrono::steady_clock::now();
  check_layout(sc, params_in);
  lw_params p = params_in;
  p.frame[2] = 0;
  const uint32_t w = p.frame[0], h = p.frame[1];
  if (w == 0 || h == 0) throw InvalidScene{"empty frame"};
  if (p.frame[3] > kMaxSuffix) throw InvalidScene{"the adjoint takes at most 16 bounces"};
  const uint64_t npix = (uint64_t)w * h;
  if (npix > (1ull << 31) / 3) throw InvalidScene{"frame too large"};
  const uint32_t spp = std::max(spp_in, 1u);
  DiffDev D = prepare_diff(b, sc, p, dd);
  const uint32_t slots = 3u * D.n_pal + D.n_int + 3u;
  b->diff_dldi.upload(dldi, npix * 3, b->stream);
  b->diff_grads.ensure(slots);
  D.dldi = b->diff_dldi.ptr;
  D.grads = b->diff_grads.ptr;
  k_zero<<<dim3((slots + kBlock - 1) / kBlock), dim3(kBlock), 0, b->stream>>>(D.grads, slots);
  HIP_TRY(hipGetLastError());
  const DevScene ds = sc.dev();
  const float inv_spp = 1.0f / (float)spp;
  const size_t lds = (size_t)slots * sizeof(double);
  lw_timing tm{};
  for (uint32_t done = 0; done < spp;) {
    const uint32_t nb = std::min(spp - done, kMaxBatch);
    HIP_TRY(hipEventRecord(b->ev[0], b->stream));
    const uint32_t lay = launch_layout(nb);
    const dim3 grid = launch_grid(lay, w, h, nb);
    LW_LAYOUT_DISPATCH(lay, LW_DIFF_DISPATCH(b->waves, (k_adjoint<W, L><<<grid, dim3(kBlock), lds, b->stream>>>(
                                                            p, ds, D, done, inv_spp))));
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(b->ev[1], b->stream));
    HIP_TRY(hipEventSynchronize(b->ev[1]));
    float ms = 0.0f;
    HIP_TRY(hipEventElapsedTime(&ms, b->ev[0], b->ev[1]));
    tm.trace_ms += ms;
    tm.launches += 1;
    done += nb;
  }
  HIP_TRY(hipMemcpyAsync(grads, D.grads, slots * sizeof(double), hipMemcpyDeviceToHost, b->stream));
  HIP_TRY(hipStreamSynchronize(b->stream));
  tm.samples = npix * spp;
  tm.wall_ms =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_wall).count();
  b->timing = tm;
}

// The adjoint of the last batch render_batch planned (b->plan), with D.dldi holding every
// view's dL/dI (laid out by pix_prefix) and D.grads the device gradient slots, which it
// zeroes first.
void adjoint_batch(lw_hip_backend *b, const lw_hip_scene &sc, const lw_params &p0,
                   const DiffDev &D, lw_timing &tm) {
  if (p0.frame[3] > kMaxSuffix) throw InvalidScene{"the adjoint takes at most 16 bounces"};
  const BatchPlan &plan = b->plan;
  const uint32_t slots = 3u * D.n_pal + D.n_int + 3u;
  const DevScene ds = sc.dev();
  HIP_TRY(hipMemcpyAsync(b->batch_dev.ptr, &plan.bd(), sizeof(BatchDev), hipMemcpyHostToDevice,
                         b->stream));
  k_zero<<<dim3((slots + kBlock - 1) / kBlock), dim3(kBlock), 0, b->stream>>>(D.grads, slots);
  HIP_TRY(hipGetLastError());
  const uint32_t blocks = (uint32_t)std::max<uint64_t>(1, plan.items / kBlock);
  const size_t lds = (size_t)slots * sizeof(double);
  HIP_TRY(hipEventRecord(b->ev[2], b->stream));
  for (uint32_t k = 0; k < plan.npass; k++) {
    k_batch_plan<<<dim3(1), dim3(kMaxViews), 0, b->stream>>>(b->batch_dev.ptr);
    if (batch_by64(plan.bd())) {
      LW_DIFF_DISPATCH(b->waves, (k_adjoint_batch<W, 2><<<dim3(blocks), dim3(kBlock), lds, b->stream>>>(
                                     p0, ds, D, b->batch_dev.ptr)));
    } else {
      LW_DIFF_DISPATCH(b->waves, (k_adjoint_batch<W, 0><<<dim3(blocks), dim3(kBlock), lds, b->stream>>>(
                                     p0, ds, D, b->batch_dev.ptr)));
    }
    HIP_TRY(hipGetLastError());
  }
  HIP_TRY(hipEventRecord(b->ev[3], b->stream));
  HIP_TRY(hipEventSynchronize(b->ev[3]));
  float ms = 0.0f;
  HIP_TRY(hipEventElapsedTime(&ms, b->ev[2], b->ev[3]));
  tm.trace_ms = ms;
  tm.launches = plan.npass;
  tm.samples = plan.samples;
}

// The ToneDev of an operator pair (post.rs:148-215).
ToneDev make_tone(float exposure, uint32_t op_a, uint32_t op_b, float weight) {
  if (op_a > LW_TONEMAP_FILMIC_MID_GREY || op_b > LW_TONEMAP_FILMIC_MID_GREY)
    throw InvalidScene{"unknown tone-mapping operator"};
  ToneDev t{};
  lw_encode_thresholds(t.thr);
  float k[8];
  lw_hable_curve(op_a, k);
  std::memcpy(t.ca, k, sizeof(t.ca));
  t.white_a = (double)k[7];
  lw_hable_curve(op_b, k);
  std::memcpy(t.cb, k, sizeof(t.cb));
  t.white_b = (double)k[7];
  t.exposure = exposure;
  t.weight = weight;
  t.op_a = op_a;
  t.op_b = op_b;
  return t;
}

// The table pow_libm or log2_libm reads: every f32 of the function's domain ([0, 1] for the
// powers, [1, 65536] for log2) whose rounded f64 value lies in the band (k_pow_band, one pass
// over the domain's bit patterns), then libm on those inputs here, keeping the ones where it
// differs from the device's rounding.
const lw_hip_backend::PowFix &ensure_pow_fix(lw_hip_backend *b, int which) {
  lw_hip_backend::PowFix &pf = b->pow_fix[which];
  if (pf.ready) return pf;
  const auto t0 = std::chrono::steady_clock::now();
  const float y = which == LIBM_LOG2 ? 0.0f : kPowExponent[which];
  const uint32_t u0 = which == LIBM_LOG2 ? 0x3f800000u : 0u;     // 1.0f or 0.0f
  const uint32_t n = which == LIBM_LOG2 ? 0x08000001u : 0x3f800001u;  // through 65536.0f or 1.0f
  const uint32_t cap = 1u << 24;  // at most about 6.0M inputs are in the band
  DevBuf<uint2> cand;
  DevBuf<uint32_t> count;
  cand.alloc(cap);
  count.alloc(1);
  HIP_TRY(hipMemsetAsync(count.ptr, 0, sizeof(uint32_t), b->stream));
  k_pow_band<<<dim3(8192), dim3(kBlock), 0, b->stream>>>(u0, n, which, y, cand.ptr, count.ptr, cap);
  HIP_TRY(hipGetLastError());
  uint32_t found = 0;
  HIP_TRY(hipMemcpyAsync(&found, count.ptr, sizeof(found), hipMemcpyDeviceToHost, b->stream));
  HIP_TRY(hipStreamSynchronize(b->stream));
  if (found > cap) throw InvalidScene{"libm band holds more inputs than expected"};
  std::vector<uint2> host(found);
  if (found)
    HIP_TRY(hipMemcpy(host.data(), cand.ptr, (size_t)found * sizeof(uint2), hipMemcpyDeviceToHost));
  std::vector<uint2> fix;
  for (const uint2 &e : host) {
    float x;
    std::memcpy(&x, &e.x, 4);
    const float g = which == LIBM_LOG2 ? lw::libm_log2f(x) : lw::libm_powf(x, y);
    uint32_t gb;
    std::memcpy(&gb, &g, 4);
    if (gb != e.y) fix.push_back(make_uint2(e.x, gb));
  }
  std::sort(fix.begin(), fix.end(), [](const uint2 &a, const uint2 &c) { return a.x < c.x; });
  pf.fix.upload(fix.data(), fix.size(), b->stream);
  HIP_TRY(hipStreamSynchronize(b->stream));
  pf.n = (uint32_t)fix.size();
  pf.ready = true;
  pf.ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return pf;
}

// Tone-maps n pixels already on the device (in) into b->tone_out and copies them to out.
void tonemap_device(lw_hip_backend *b, const float *in, size_t n, const ToneDev &t, bool mix,
                    uint32_t *out) {
  if (n > 0xFFFFFFFFull) throw InvalidScene{"image too large"};
  b->tone_out.ensure(std::max<size_t>(n, 1));
  const dim3 grid((unsigned)((n + kBlock - 1) / kBlock));
  HIP_TRY(hipEventRecord(b->ev[0], b->stream));
  if (n) {
    if (mix)
      k_tonemap_mix<<<grid, dim3(kBlock), 0, b->stream>>>(in, (uint32_t)n, t, b->pow_fix[POW_ENCODE].fix.ptr,
                                                          b->pow_fix[POW_ENCODE].n, b->tone_out.ptr);
    else
      k_tonemap<<<grid, dim3(kBlock), 0, b->stream>>>(in, (uint32_t)n, t, b->tone_out.ptr);
    HIP_TRY(hipGetLastError());
  }
  HIP_TRY(hipEventRecord(b->ev[1], b->stream));
  if (n)
    HIP_TRY(hipMemcpyAsync(out, b->tone_out.ptr, n * sizeof(uint32_t), hipMemcpyDeviceToHost, b->stream));
  HIP_TRY(hipStreamSynchronize(b->stream));
  float ms = 0.0f;
  HIP_TRY(hipEventElapsedTime(&ms, b->ev[0], b->ev[1]));
  lw_timing tm{};
  tm.trace_ms = ms;
  tm.launches = 1;
  tm.samples = n;
  b->timing = tm;
}

template <typename F>
int guarded(lw_hip_backend *b, F &&f) {
  try {
    f();
    if (b) b->last_error.clear();
    return LW_OK;
  } catch (const InvalidScene &e) {
    set_err(b, "invalid scene: " + e.what);
    return LW_EINVAL;
  } catch (const Unsupported &e) {
    set_err(b, "unsupported: " + e.what);
    return LW_EUNSUPPORTED;
  } catch (const HipError &e) {
    set_err(b, "HIP: " + e.what);
    return e.code == hipErrorOutOfMemory ? LW_ENOMEM : LW_EHIP;
  } catch (const std::exception &e) {
    set_err(b, e.what());
    return LW_EHIP;
  }
}

}  // namespace

// An inverse fit's state on the device (lw_hip_fit_*): the views, the parameter groups, the
// factors with their Adam moments, the targets, dL/dI and the gradient.
struct lw_hip_fit {
  std::vector<lw_view> views;
  uint32_t n_pal = 0, n_int = 0, t = 0;
  lw_fit_options opt{};
  DevBuf<float> x, m, v, target, dldi;
  DevBuf<uint32_t> qg, qig;
  DevBuf<double> grads, loss;
  DevBuf<float> dummy;  // the derivative table a fit without sky parameters points at
  uint64_t npix_total = 0;
};

extern "C" {

const char *lw_last_error(void) { return g_last_error.c_str(); }

int lw_hip_probe(char *desc, size_t cap) {
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  std::string s;
  int rc = LW_OK;
  if (e != hipSuccess || n == 0) {
    s = std::string("no HIP device: ") + (e != hipSuccess ? hipGetErrorString(e) : "device count 0");
    rc = LW_ENODEV;
    g_last_error = s;
  } else {
    hipDeviceProp_t prop;
    e = hipGetDeviceProperties(&prop, 0);
    if (e != hipSuccess) {
      s = std::string("hipGetDeviceProperties: ") + hipGetErrorString(e);
      rc = LW_EHIP;
      g_last_error = s;
    } else {
      s = std::string(prop.name) + " (" + prop.gcnArchName + ", " +
          std::to_string(prop.multiProcessorCount) + " CUs)";
    }
  }
  if (desc && cap) {
    std::strncpy(desc, s.c_str(), cap - 1);
    desc[cap - 1] = '\0';
  }
  return rc;
}

int lw_hip_open(int device, lw_hip_backend **out) {
  if (!out) return LW_EINVAL;
  *out = nullptr;
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n == 0) {
    g_last_error = std::string("no HIP device: ") +
                   (e != hipSuccess ? hipGetErrorString(e) : "device count 0");
    return LW_ENODEV;
  }
  if (device < 0 || device >= n) {
    g_last_error = "device index out of range";
    return LW_EINVAL;
  }
  auto *b = new lw_hip_backend();
  b->device = device;
  int rc = guarded(b, [&] {
    HIP_TRY(hipSetDevice(device));
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, device));
    std::string nm = std::string(prop.name) + " (hip)";
    for (auto &c : nm) c = (char)std::tolower((unsigned char)c);
    b->name = nm;
    HIP_TRY(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    for (auto &ev : b->ev) HIP_TRY(hipEventCreate(&ev));
    b->counter.alloc(1);
    if (device < kMaxScratchDevices) {  // once per device, with the device current
      std::lock_guard<std::mutex> lock(g_scratch_mu);
      if (!g_scratch_ptr[device]) {
        void *ptr = nullptr;
        HIP_TRY(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        g_scratch_ptr[device] = ptr;
      }
      b->scratch = g_scratch_ptr[device];
    }
    if (const char *e = std::getenv("LW_WAVES")) b->waves = b->tiles_waves = std::atoi(e);
    if (const char *e = std::getenv("LW_TILES_WAVES")) b->tiles_waves = std::atoi(e);
    if (const char *e = std::getenv("LW_TRACE_WAVES")) b->trace_waves = std::atoi(e);
    int per_cu_frame = 0, per_cu_list = 0;
    LW_WAVES_DISPATCH(b->waves, {
      HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(
          &per_cu_frame, reinterpret_cast<const void *>(&k_paths<FrameWork, W>), kBlock, 0));
      HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(
          &per_cu_list, reinterpret_cast<const void *>(&k_paths<ListWork, W>), kBlock, 0));
    });
    const uint32_t cus = (uint32_t)std::max(1, prop.multiProcessorCount);
    b->blocks_frame = cus * (uint32_t)std::max(1, per_cu_frame);
    b->blocks_list = cus * (uint32_t)std::max(1, per_cu_list);
    if (const char *e = std::getenv("LW_PERSISTENT_BLOCKS_PER_CU")) {
      const int v = std::atoi(e);
      if (v > 0) b->blocks_frame = b->blocks_list = cus * (uint32_t)v;
    }
    if (const char *e = std::getenv("LW_KERNEL")) {
      if (std::strcmp(e, "direct") == 0) b->mode = MODE_DIRECT;
      if (std::strcmp(e, "persistent") == 0) b->mode = MODE_PERSISTENT;
      if (std::strcmp(e, "wavefront") == 0) b->mode = MODE_WAVEFRONT;
    }
    if (const char *e = std::getenv("LW_XCD_REMAP")) b->xcd_remap = std::atoi(e) != 0;
    if (const char *e = std::getenv("LW_SPL")) b->per_lane = (uint32_t)std::max(1, std::min(64, std::atoi(e)));
    if (const char *e = std::getenv("LW_WF_SLOTS")) {
      const long v = std::atol(e);
      if (v >= kBlock) b->wf_slots = (uint32_t)std::min<long>(v, 1l << 24) / kBlock * kBlock;
    }
    HIP_TRY(hipHostMalloc((void **)&b->pinned, 64, hipHostMallocDefault));
    HIP_TRY(hipHostMalloc((void **)&b->progress_host, 64, hipHostMallocMapped | hipHostMallocCoherent));
    HIP_TRY(hipHostGetDevicePointer((void **)&b->progress_dev, b->progress_host, 0));
    if (const char *e = std::getenv("LW_FRAME_FORM")) b->frame_launches = std::strcmp(e, "launches") == 0;
    if (const char *e = std::getenv("LW_BATCH_GRAPHS")) b->use_graphs = std::strcmp(e, "0") != 0;
    if (const char *e = std::getenv("LW_BATCH_PASSES")) b->batch_passes = (uint32_t)std::max(0, std::atoi(e));
    if (const char *e = std::getenv("LW_BATCH_FORM")) b->batch_passes_form = std::strcmp(e, "passes") == 0;
    set_tiles_waves(b, b->tiles_waves);
  });
  if (rc != LW_OK) {
    lw_hip_close(b);
    return rc;
  }
  *out = b;
  return LW_OK;
}

const char *lw_hip_name(const lw_hip_backend *b) { return b ? b->name.c_str() : ""; }

const char *lw_hip_last_error(const lw_hip_backend *b) { return b ? b->last_error.c_str() : ""; }

int lw_hip_render_quick(lw_hip_backend *b, const lw_quick_frame *f, uint32_t *out) {
  if (!b || !f || !out || !f->sky) return LW_EINVAL;
  return guarded(b, [&] {
    auto fail = [](const std::string &m) { throw InvalidScene{m}; };
    if (f->width == 0 || f->height == 0 || (uint64_t)f->width * f->height >= (1ull << 31))
      fail("frame size out of range");
    // Whole triangles only: a trailing partial triangle is dropped, as raster.rs:199 drops it.
    const size_t n_opaque = f->opaque_len / 3 * 3, n_translucent = f->translucent_len / 3 * 3,
                 n_see_through = f->see_through_len / 3 * 3;
    const size_t n_idx = n_opaque + n_translucent + n_see_through;
    if ((n_idx && !f->vertices) || f->vertices_len > 0xFFFFFFFFull || n_idx / 3 > 0x3FFFFFFFull) fail("vertices");
    std::vector<uint32_t> idx;
    idx.reserve(n_idx);
    for (auto list : {std::make_pair(f->opaque, n_opaque), std::make_pair(f->translucent, n_translucent),
                      std::make_pair(f->see_through, n_see_through)}) {
      if (list.second && !list.first) fail("an index list is null");
      for (size_t k = 0; k < list.second; k++) {
        if (list.first[k] >= f->vertices_len) fail("an index names a vertex past the vertex list");
        idx.push_back(list.first[k]);
      }
    }
    const uint32_t aw = f->atlas_size[0], ah = f->atlas_size[1];
    if (!f->atlas_pixels || aw == 0 || ah == 0 || (uint64_t)aw * ah > 0xFFFFFFFFull) fail("atlas");
    if (f->mip_levels > lwq::kMaxMipLevels || (f->mip_levels && ((aw >> f->mip_levels) == 0 || (ah >> f->mip_levels) == 0)))
      fail("mip levels out of range");
    lwq::Input in{};
    uint64_t mip_words = 0;
    for (uint32_t l = 1; l <= f->mip_levels; l++) {
      in.mip_offset[l] = (uint32_t)mip_words;
      mip_words += (uint64_t)(aw >> l) * (ah >> l);
    }
    if (mip_words > f->atlas_mips_len || (mip_words && !f->atlas_mips)) fail("atlas mips shorter than their levels");
    const size_t nbins = (size_t)LW_STAR_BINS_AZ * LW_STAR_BINS_EL;
    if (!f->stars || f->stars_len < nbins + 1) fail("star table shorter than its bins");
    for (size_t k = 0; k < nbins; k++)
      if (f->stars[k] > f->stars[k + 1]) fail("star table bins out of order");
    if (nbins + 1 + (size_t)f->stars[nbins] * LW_STAR_WORDS > f->stars_len) fail("star table entries past its end");
    HIP_TRY(hipSetDevice(b->device));
    const auto &fix = ensure_pow_fix(b, LIBM_LOG2);
    if (!b->quick_ws) b->quick_ws = lwq::workspace_new();
    auto put = [&](auto &buf, const auto *src, size_t count) {
      buf.ensure(std::max<size_t>(count, 1));
      if (count) HIP_TRY(hipMemcpyAsync(buf.ptr, src, count * sizeof(*src), hipMemcpyHostToDevice, b->stream));
      return buf.ptr;
    };
    auto &dverts = b->quick_verts;
    auto &didx = b->quick_idx, &datlas = b->quick_atlas, &dmips = b->quick_mips, &dstars = b->quick_stars;
    auto &dout = b->quick_out;
    put(dverts, f->vertices, f->vertices_len);
    put(didx, idx.data(), idx.size());
    put(datlas, f->atlas_pixels, (size_t)aw * ah);
    put(dmips, f->atlas_mips, (size_t)mip_words);
    put(dstars, f->stars, f->stars_len);
    dout.ensure((size_t)f->width * f->height);
    in.verts = dverts.ptr;
    in.n_verts = (uint32_t)f->vertices_len;
    in.indices = didx.ptr;
    in.n_opaque = (uint32_t)(n_opaque / 3);
    in.n_translucent = (uint32_t)(n_translucent / 3);
    in.n_see_through = (uint32_t)(n_see_through / 3);
    in.atlas = datlas.ptr;
    in.atlas_w = aw;
    in.atlas_h = ah;
    in.mips = dmips.ptr;
    in.levels = f->mip_levels;
    in.mip_rows = f->mip_rows;
    in.stars = dstars.ptr;
    in.width = f->width;
    in.height = f->height;
    std::memcpy(in.view_proj, f->view_proj, sizeof(in.view_proj));
    in.sky = *f->sky;
    lw_disc_sector_trig(in.sectors.s0, in.sectors.c0, in.sectors.s1, in.sectors.c1);
    in.log2_fix = fix.fix.ptr;
    in.n_log2_fix = fix.n;
    lwq::Stats st;
    try {
      lwq::render(b->stream, in, *b->quick_ws, dout.ptr, st);
    } catch (const lwq::Error &e) {
      if (e.invalid) throw InvalidScene{e.what};
      throw HipError{e.code, e.what};
    }
    HIP_TRY(hipMemcpy(out, dout.ptr, (size_t)f->width * f->height * sizeof(uint32_t), hipMemcpyDeviceToHost));
    lw_timing tm{};
    tm.trace_ms = st.kernel_ms;
    tm.launches = 1;
    tm.samples = (uint64_t)f->width * f->height;
    b->timing = tm;
  });
}

void lw_hip_close(lw_hip_backend *b) {
  if (!b) return;
  (void)hipSetDevice(b->device);
  b->rad.free();
  b->acc.free();
  b->out.free();
  b->wf_st.free();
  b->wf_o.free();
  b->wf_d.free();
  b->wf_hit.free();
  b->wf_queue.free();
  b->wf_ctr.free();
  b->batch_dev.free();
  if (b->quick_ws) lwq::workspace_free(b->quick_ws);
  if (b->pinned) (void)hipHostFree(b->pinned);
  if (b->progress_host) (void)hipHostFree(b->progress_host);
  for (auto &ev : b->ev)
    if (ev) (void)hipEventDestroy(ev);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

int lw_hip_scene_upload(lw_hip_backend *b, const lw_scene_view *view, lw_hip_scene **out) {
  if (!b || !view || !out) return LW_EINVAL;
  *out = nullptr;
  auto *sc = new lw_hip_scene();
  int rc = guarded(b, [&] { upload_scene(b, *view, *sc); });
  if (rc != LW_OK) {
    delete sc;
    return rc;
  }
  *out = sc;
  return LW_OK;
}

void lw_hip_scene_free(lw_hip_backend *b, lw_hip_scene *scene) {
  if (b) (void)hipSetDevice(b->device);
  delete scene;
}

int lw_hip_scene_build(lw_hip_backend *b, const lw_voxel_scene *vs, lw_scene_info *info,
                       lw_build_stats *stats, lw_hip_scene **out) {
  if (!b || !vs || !info || !out) return LW_EINVAL;
  *out = nullptr;
  auto *sc = new lw_hip_scene();
  int rc = guarded(b, [&] { build_scene(b, *vs, *sc, *info, stats); });
  if (rc != LW_OK) {
    delete sc;
    return rc;
  }
  *out = sc;
  return LW_OK;
}

int lw_hip_scene_read(lw_hip_backend *b, const lw_hip_scene *scene, uint32_t which, void *dst,
                      size_t cap, size_t *len) {
  if (!b || !scene || !len) return LW_EINVAL;
  return guarded(b, [&] {
    const void *src = nullptr;
    size_t n = 0, size = 0;
    switch (which) {
      case LW_SCENE_GRID: src = scene->grid.ptr; n = scene->grid.n; size = 4; break;
      case LW_SCENE_EMITTERS: src = scene->emitters.ptr; n = scene->emitters.n; size = sizeof(lw_emitter); break;
      case LW_SCENE_EMITTER_GRID: src = scene->egrid.ptr; n = scene->egrid.n; size = 4; break;
      case LW_SCENE_QUADS: src = scene->quads.ptr; n = scene->quads.n; size = sizeof(lw_quad); break;
      case LW_SCENE_OCTANT_BOXES: src = scene->octant_boxes.ptr; n = scene->octant_boxes.n; size = 1; break;
      case LW_SCENE_CELL_CUBES: src = scene->cell_ext.ptr; n = scene->cell_ext.n; size = 4; break;
      default: throw InvalidScene{"unknown scene buffer"};
    }
    *len = n;
    if (!dst) return;
    if (cap < n) throw InvalidScene{"destination smaller than the buffer"};
    HIP_TRY(hipSetDevice(b->device));
    if (n) HIP_TRY(hipMemcpy(dst, src, n * size, hipMemcpyDeviceToHost));
  });
}

int lw_hip_render_hq_scene(lw_hip_backend *b, const lw_hip_scene *scene, const lw_params *params,
                           uint32_t spp, lw_progress_fn progress, void *user, float *out_rgb) {
  if (!b || !scene || !params || !out_rgb) return LW_EINVAL;
  return guarded(b, [&] { render_frame(b, *scene, *params, spp, progress, user, out_rgb); });
}

int lw_hip_render_hq_scene_noalloc(lw_hip_backend *b, const lw_hip_scene *scene,
                                   const lw_params *params, uint32_t spp,
                                   lw_progress_fn progress, void *user, float *out_rgb) {
  if (!b || !scene || !params || !out_rgb) return LW_EINVAL;
  return guarded(b, [&] {
    if (!b->scratch) throw Unsupported{"no scratch memory on this device"};
    const uint64_t npix = (uint64_t)params->frame[0] * params->frame[1];
    FrameScratch fs;
    if (!split_scratch(b->scratch, npix, fs))
      throw InvalidScene{"a frame of " + std::to_string(npix) + " pixels needs " +
                         std::to_string((npix * 24 + 255) / 256 * 256 + npix * 12) +
                         " bytes of scratch memory; it has " + std::to_string(kScratchBytes)};
    render_frame(b, *scene, *params, spp, progress, user, out_rgb, nullptr, nullptr, false, &fs);
  });
}

int lw_hip_count_traces(lw_hip_backend *b, const lw_hip_scene *scene, const lw_params *params, uint32_t spp,
                        uint64_t *traces) {
  if (!b || !scene || !params || !traces) return LW_EINVAL;
  return guarded(b, [&] {
    check_layout(*scene, *params);
    lw_params p = *params;
    p.frame[2] = 0;
    const uint32_t w = p.frame[0], h = p.frame[1];
    if (w == 0 || h == 0) throw InvalidScene{"empty frame"};
    const uint32_t tiles = ((w + 15) / 16) * ((h + 15) / 16);
    HIP_TRY(hipSetDevice(b->device));
    const DevScene ds = scene->dev();
    DevBuf<unsigned long long> total;
    total.alloc(1);
    HIP_TRY(hipMemsetAsync(total.ptr, 0, sizeof(unsigned long long), b->stream));
    const uint32_t n = std::max(spp, 1u);
    for (uint32_t done = 0; done < n; done += 4096u) {
      const uint32_t nb = std::min(4096u, n - done);
      k_count_traces<<<dim3(tiles, nb), dim3(kBlock), 0, b->stream>>>(p, ds, done, total.ptr);
      HIP_TRY(hipGetLastError());
    }
    unsigned long long t = 0;
    HIP_TRY(hipMemcpyAsync(&t, total.ptr, sizeof(t), hipMemcpyDeviceToHost, b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    *traces = t;
  });
}

int lw_hip_render_hq(lw_hip_backend *b, const lw_scene_view *view, uint32_t spp,
                     lw_progress_fn progress, void *user, float *out_rgb) {
  if (!b || !view || !out_rgb) return LW_EINVAL;
  return guarded(b, [&] {
    lw_hip_scene sc;
    upload_scene(b, *view, sc);
    render_frame(b, sc, *view->params, spp, progress, user, out_rgb);
  });
}

int lw_hip_render_hq_batch(lw_hip_backend *b, const lw_hip_scene *scene, const lw_view *views,
                           size_t n, lw_progress_fn progress, void *user, float *const *out_rgb) {
  if (!b || !scene || !views || !out_rgb) return LW_EINVAL;
  return guarded(b, [&] { render_batch(b, *scene, views, n, progress, user, out_rgb); });
}

int lw_hip_render_diff(lw_hip_backend *b, const lw_hip_scene *scene, const lw_params *params,
                       uint32_t spp, const lw_diff_desc *diff, float *out_rgb, double *out_f64) {
  if (!b || !scene || !params || !diff || !out_rgb) return LW_EINVAL;
  return guarded(b, [&] {
    const DiffDev D = prepare_diff(b, *scene, *params, *diff);
    // No progress to report: launches of up to kMaxBatch samples per pixel, as the adjoint's.
    render_frame(b, *scene, *params, spp, nullptr, nullptr, out_rgb, &D, out_f64, true);
  });
}

int lw_hip_adjoint(lw_hip_backend *b, const lw_hip_scene *scene, const lw_params *params,
                   uint32_t spp, const lw_diff_desc *diff, const float *d_loss_d_image,
                   double *grads) {
  if (!b || !scene || !params || !diff || !d_loss_d_image || !grads) return LW_EINVAL;
  return guarded(b, [&] { adjoint_frame(b, *scene, *params, spp, *diff, d_loss_d_image, grads); });
}

int lw_hip_render_diff_batch(lw_hip_backend *b, const lw_hip_scene *scene, const lw_view *views,
                             size_t n, const lw_diff_desc *diff, float *const *out_rgb) {
  if (!b || !scene || !views || !diff || !out_rgb) return LW_EINVAL;
  return guarded(b, [&] {
    if (n == 0) throw InvalidScene{"a batch holds 1 to 256 views"};
    const DiffDev D = prepare_diff(b, *scene, views[0].params, *diff);
    render_batch(b, *scene, views, n, nullptr, nullptr, out_rgb, &D);
  });
}

int lw_hip_adjoint_batch(lw_hip_backend *b, const lw_hip_scene *scene, const lw_view *views,
                         size_t n, const lw_diff_desc *diff, const float *const *d_loss_d_image,
                         double *grads) {
  if (!b || !scene || !views || !diff || !d_loss_d_image || !grads) return LW_EINVAL;
  return guarded(b, [&] {
    if (n == 0) throw InvalidScene{"a batch holds 1 to 256 views"};
    DiffDev D = prepare_diff(b, *scene, views[0].params, *diff);
    b->plan = plan_batch(b, *scene, views, n);
    const BatchDev &bd = b->plan.bd();
    const uint32_t slots = 3u * D.n_pal + D.n_int + 3u;
    b->diff_dldi.ensure(b->plan.npix_total * 3);
    for (size_t v = 0; v < n; v++) {
      if (!d_loss_d_image[v]) throw InvalidScene{"null dL/dI for a view"};
      HIP_TRY(hipMemcpyAsync(b->diff_dldi.ptr + (size_t)bd.pix_prefix[v] * 3, d_loss_d_image[v],
                             (size_t)bd.cam[v].width * bd.cam[v].height * 3 * sizeof(float),
                             hipMemcpyHostToDevice, b->stream));
    }
    b->diff_grads.ensure(slots);
    b->batch_dev.ensure(1);
    D.dldi = b->diff_dldi.ptr;
    D.grads = b->diff_grads.ptr;
    lw_timing tm{};
    adjoint_batch(b, *scene, views[0].params, D, tm);
    HIP_TRY(hipMemcpyAsync(grads, D.grads, slots * sizeof(double), hipMemcpyDeviceToHost, b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    b->timing = tm;
  });
}

int lw_hip_fit_new(lw_hip_backend *b, const lw_hip_scene *scene, const lw_view *views, size_t n,
                   const lw_diff_desc *diff, const float *const *targets,
                   const lw_fit_options *opt, lw_hip_fit **out) {
  if (!b || !scene || !views || !diff || !targets || !opt || !out) return LW_EINVAL;
  *out = nullptr;
  auto *f = new (std::nothrow) lw_hip_fit();
  if (!f) return LW_ENOMEM;
  const int rc = guarded(b, [&] {
    if (n == 0) throw InvalidScene{"a batch holds 1 to 256 views"};
    if (diff->sky_params) throw Unsupported{"an inverse fit over the sky parameters"};
    if (!(opt->lr > 0.0f) || !(opt->beta1 >= 0.0f && opt->beta1 < 1.0f) ||
        !(opt->beta2 >= 0.0f && opt->beta2 < 1.0f) || !(opt->eps > 0.0f) || !(opt->min_factor > 0.0f))
      throw InvalidScene{"bad Adam options"};
    if (views[0].params.frame[3] > kMaxSuffix) throw InvalidScene{"the adjoint takes at most 16 bounces"};
    (void)prepare_diff(b, *scene, views[0].params, *diff);  // checks the groups and factors
    BatchPlan plan = plan_batch(b, *scene, views, n);
    const BatchDev &bd = plan.bd();
    f->views.assign(views, views + n);
    f->n_pal = diff->n_palette;
    f->n_int = diff->n_intensity;
    f->opt = *opt;
    f->npix_total = plan.npix_total;
    const uint32_t np = 3 * f->n_pal + f->n_int;
    std::vector<float> x0(std::max(np, 1u), 0.0f);
    for (uint32_t i = 0; i < 3 * f->n_pal; i++) x0[i] = diff->palette[i];
    for (uint32_t i = 0; i < f->n_int; i++) x0[3 * f->n_pal + i] = diff->intensity[i];
    f->x.upload(x0.data(), x0.size(), b->stream);
    std::vector<float> zeros(x0.size(), 0.0f);
    f->m.upload(zeros.data(), zeros.size(), b->stream);
    f->v.upload(zeros.data(), zeros.size(), b->stream);
    const size_t nq = scene->quad_flags.size();
    f->qg.upload(diff->quad_palette, nq, b->stream);
    f->qig.upload(diff->quad_intensity, nq, b->stream);
    f->target.alloc(plan.npix_total * 3);
    f->dldi.alloc(plan.npix_total * 3);
    for (size_t v = 0; v < n; v++) {
      if (!targets[v]) throw InvalidScene{"null target for a view"};
      HIP_TRY(hipMemcpyAsync(f->target.ptr + (size_t)bd.pix_prefix[v] * 3, targets[v],
                             (size_t)bd.cam[v].width * bd.cam[v].height * 3 * sizeof(float),
                             hipMemcpyHostToDevice, b->stream));
    }
    f->grads.alloc(np + 3);
    f->loss.alloc(1);
    f->dummy.alloc(1);
    HIP_TRY(hipStreamSynchronize(b->stream));
  });
  if (rc != LW_OK) {
    delete f;
    return rc;
  }
  *out = f;
  return LW_OK;
}

int lw_hip_fit_step(lw_hip_backend *b, const lw_hip_scene *scene, lw_hip_fit *f,
                    lw_fit_stats *stats) {
  if (!b || !scene || !f || !stats) return LW_EINVAL;
  return guarded(b, [&] {
    auto t_wall = std::chrono::steady_clock::now();
    DiffDev D{};
    D.pal = f->x.ptr;
    D.inten = f->x.ptr + 3 * f->n_pal;
    D.qgroup = f->qg.ptr;
    D.qigroup = f->qig.ptr;
    D.dsky = f->dummy.ptr;
    D.dldi = f->dldi.ptr;
    D.grads = f->grads.ptr;
    D.n_pal = f->n_pal;
    D.n_int = f->n_int;
    const lw_view *views = f->views.data();
    const size_t n = f->views.size();
    render_batch(b, *scene, views, n, nullptr, nullptr, nullptr, &D);
    const double fwd_ms = b->timing.trace_ms;
    k_zero<<<dim3(1), dim3(kBlock), 0, b->stream>>>(f->loss.ptr, 1);
    const uint64_t n3 = f->npix_total * 3;
    k_fit_loss<<<dim3((unsigned)((n3 + kBlock - 1) / kBlock)), dim3(kBlock), 0, b->stream>>>(
        b->batch_dev.ptr, b->acc.ptr, f->target.ptr, f->dldi.ptr, 1.0 / (double)n3, f->loss.ptr);
    HIP_TRY(hipGetLastError());
    lw_timing ta{};
    adjoint_batch(b, *scene, views[0].params, D, ta);
    f->t += 1;
    const float bc1 = 1.0f - std::pow(f->opt.beta1, (float)f->t);
    const float bc2 = 1.0f - std::pow(f->opt.beta2, (float)f->t);
    const uint32_t np = 3 * f->n_pal + f->n_int;
    if (np > 0) {
      k_adam<<<dim3((np + kBlock - 1) / kBlock), dim3(kBlock), 0, b->stream>>>(
          f->x.ptr, f->grads.ptr, f->m.ptr, f->v.ptr, np, f->opt.lr, f->opt.beta1, f->opt.beta2,
          f->opt.eps, bc1, bc2, f->opt.min_factor);
      HIP_TRY(hipGetLastError());
    }
    double loss = 0.0;
    HIP_TRY(hipMemcpyAsync(&loss, f->loss.ptr, sizeof(double), hipMemcpyDeviceToHost, b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    stats->loss = loss;
    stats->forward_ms = fwd_ms;
    stats->adjoint_ms = ta.trace_ms;
    stats->wall_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_wall).count();
    stats->step = f->t;
    stats->samples = b->plan.samples;
  });
}

int lw_hip_fit_params(lw_hip_backend *b, const lw_hip_fit *f, float *palette, float *intensity,
                      double *grads) {
  if (!b || !f) return LW_EINVAL;
  return guarded(b, [&] {
    const uint32_t np = 3 * f->n_pal + f->n_int;
    std::vector<float> x(std::max(np, 1u));
    HIP_TRY(hipMemcpyAsync(x.data(), f->x.ptr, x.size() * sizeof(float), hipMemcpyDeviceToHost,
                           b->stream));
    if (grads)
      HIP_TRY(hipMemcpyAsync(grads, f->grads.ptr, (np + 3) * sizeof(double), hipMemcpyDeviceToHost,
                             b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    if (palette) std::memcpy(palette, x.data(), 3 * f->n_pal * sizeof(float));
    if (intensity) std::memcpy(intensity, x.data() + 3 * f->n_pal, f->n_int * sizeof(float));
  });
}

void lw_hip_fit_free(lw_hip_backend *b, lw_hip_fit *f) {
  if (!f) return;
  if (b) (void)hipSetDevice(b->device);
  delete f;
}

int lw_hip_autotune(lw_hip_backend *b, const lw_hip_scene *scene, const lw_params *params,
                    uint32_t probe_spp, lw_tune_result *out) {
  if (!b || !scene || !params || !out) return LW_EINVAL;
  const int mode0 = b->mode, waves0 = b->waves, tiles0 = b->tiles_waves;
  const uint32_t lane0 = b->per_lane;
  const int rc = guarded(b, [&] {
    struct Cand {
      int mode, waves;
      uint32_t per_lane;
    };
    static const Cand cands[] = {{MODE_DIRECT, 6, 1}, {MODE_DIRECT, 5, 1}, {MODE_DIRECT, 4, 1},
                                 {MODE_DIRECT, 8, 1}, {MODE_DIRECT, 6, 4}, {MODE_DIRECT, 5, 4},
                                 {MODE_PERSISTENT, 4, 1}, {MODE_PERSISTENT, 3, 1},
                                 {MODE_WAVEFRONT, 4, 1}};
    const uint32_t n = (uint32_t)(sizeof(cands) / sizeof(cands[0]));
    lw_params p = *params;
    const uint32_t spp = std::max(1u, std::min(probe_spp, kMaxBatch));
    std::vector<float> img((size_t)p.frame[0] * p.frame[1] * 3);
    std::memset(out, 0, sizeof(*out));
    double best = 1e300;
    for (uint32_t i = 0; i < n; i++) {
      b->mode = cands[i].mode;
      b->waves = cands[i].waves;
      b->per_lane = cands[i].per_lane;
      // warm (code object load), then time one launch of spp samples per pixel
      render_frame(b, *scene, p, 1, nullptr, nullptr, img.data());
      render_frame(b, *scene, p, spp, nullptr, nullptr, img.data(), nullptr, nullptr, true);
      const double ms = b->timing.trace_ms;
      out->entries[i] = lw_tune_entry{(uint32_t)cands[i].mode, (uint32_t)cands[i].waves,
                                      cands[i].per_lane, 0u, ms};
      if (ms < best) {
        best = ms;
        out->best = i;
      }
    }
    out->n = n;
    b->mode = cands[out->best].mode;
    b->waves = cands[out->best].waves;
    b->per_lane = cands[out->best].per_lane;
    // The tile kernel's budget, on the frame's first kTilesMinSpp samples per pixel.
    out->best_tiles = UINT32_MAX;
    static const int tile_waves[] = {4, 5, 6};
    double best_tiles = 1e300;
    p.frame[2] = 0;
    // (timed whatever launch shape won: the tile kernel serves direct-form frames)
    const int mode_won = b->mode;
    const uint32_t lane_won = b->per_lane;
    b->mode = MODE_DIRECT;
    b->per_lane = 1;
    for (int tw : tile_waves) {
      set_tiles_waves(b, tw);
      if (!uses_tiles(b, p.frame[0], p.frame[1], kTilesMinSpp)) continue;
      const auto t0 = std::chrono::steady_clock::now();
      render_frame_tiles(b, *scene, p, kTilesMinSpp, nullptr, nullptr, img.data(), t0);  // warm
      render_frame_tiles(b, *scene, p, kTilesMinSpp, nullptr, nullptr, img.data(), t0);
      const double ms = b->timing.trace_ms;
      out->entries[out->n] = lw_tune_entry{LW_FORM_TILES, (uint32_t)tw, 1u, 0u, ms};
      if (ms < best_tiles) {
        best_tiles = ms;
        out->best_tiles = out->n;
      }
      out->n++;
    }
    b->mode = mode_won;
    b->per_lane = lane_won;
    set_tiles_waves(b, out->best_tiles != UINT32_MAX ? (int)out->entries[out->best_tiles].waves : tiles0);
  });
  if (rc != LW_OK) {
    b->mode = mode0;
    b->waves = waves0;
    b->per_lane = lane0;
    try {
      set_tiles_waves(b, tiles0);
    } catch (...) {  // keep the first error's message
    }
  }
  return rc;
}

int lw_hip_tonemap(lw_hip_backend *b, const float *hdr, size_t n, float exposure, uint32_t op,
                   uint32_t *out) {
  if (!b || (n && (!hdr || !out))) return LW_EINVAL;
  return guarded(b, [&] {
    HIP_TRY(hipSetDevice(b->device));
    const ToneDev t = make_tone(exposure, op, op, 1.0f);
    b->tone_in.upload(hdr, n * 3, b->stream);
    tonemap_device(b, b->tone_in.ptr, n, t, false, out);
  });
}

int lw_hip_tonemap_mix(lw_hip_backend *b, const float *hdr, size_t n, float exposure,
                       uint32_t op_a, uint32_t op_b, float weight, uint32_t *out) {
  if (!b || (n && (!hdr || !out))) return LW_EINVAL;
  return guarded(b, [&] {
    HIP_TRY(hipSetDevice(b->device));
    // post.rs:183-188: the ends are each curve alone, bytes included
    const bool ends = weight >= 1.0f || weight <= 0.0f;
    const uint32_t op = weight >= 1.0f ? op_a : op_b;
    const ToneDev t = ends ? make_tone(exposure, op, op, 1.0f) : make_tone(exposure, op_a, op_b, weight);
    if (!ends) ensure_pow_fix(b, POW_ENCODE);
    b->tone_in.upload(hdr, n * 3, b->stream);
    tonemap_device(b, b->tone_in.ptr, n, t, !ends, out);
  });
}

int lw_hip_sky_cache(lw_hip_backend *b, const lw_params *params, float *out, size_t cap, size_t *len) {
  if (!b || !params || !len) return LW_EINVAL;
  return guarded(b, [&] {
    const float res = params->misc_f[2];
    if (!(res >= 0.0f) || res > 4096.0f) throw InvalidScene{"sky cache resolution out of range (0 .. 4096)"};
    const uint32_t n = as_u32(res);
    const size_t need = n ? (size_t)(n + 1) * (n + 1) * 3 : 0;
    *len = need;
    if (!out || !need) return;
    if (cap < need) throw InvalidScene{"destination smaller than the sky cache"};
    HIP_TRY(hipSetDevice(b->device));
    DevBuf<float> dst;
    dst.alloc(need);
    const uint32_t entries = (n + 1) * (n + 1);
    HIP_TRY(hipEventRecord(b->ev[0], b->stream));
    k_sky_cache<<<dim3((entries + kBlock - 1) / kBlock), dim3(kBlock), 0, b->stream>>>(*params, n, dst.ptr);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(b->ev[1], b->stream));
    HIP_TRY(hipMemcpyAsync(out, dst.ptr, need * sizeof(float), hipMemcpyDeviceToHost, b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    float ms = 0.0f;
    HIP_TRY(hipEventElapsedTime(&ms, b->ev[0], b->ev[1]));
    lw_timing tm{};
    tm.trace_ms = ms;
    tm.launches = 1;
    tm.samples = entries;
    b->timing = tm;
  });
}

int lw_hip_blend_chunky_sky(lw_hip_backend *b, const lw_params *params, float weight, float *maps,
                            size_t maps_len) {
  if (!b || !params || !maps) return LW_EINVAL;
  return guarded(b, [&] {
    const float *sz = params->sky_map;
    for (int k = 0; k < 4; k++)
      if (!(sz[k] >= 1.0f) || sz[k] > 65536.0f) throw InvalidScene{"sky map sizes out of range (1 .. 65536)"};
    const uint32_t lw_ = as_u32(sz[0]), lh = as_u32(sz[1]), gw = as_u32(sz[2]), gh = as_u32(sz[3]);
    const size_t lens_n = (size_t)lw_ * lh, light_n = (size_t)gw * gh;
    if (maps_len != (lens_n + light_n) * 3) throw InvalidScene{"maps_len is not the two maps' size"};
    if (lens_n > 0xFFFFFFFFull || light_n > 0xFFFFFFFFull) throw InvalidScene{"sky map too large"};
    const float res = params->misc_f[2];
    if (!(res >= 1.0f) || res > 4096.0f) throw InvalidScene{"the blend needs a sky cache resolution (1 .. 4096)"};
    const uint32_t n = as_u32(res);
    HIP_TRY(hipSetDevice(b->device));
    std::vector<float> tab(2 * ((size_t)lw_ + lh + gw + gh));
    lw_map_trig(lw_, lh, tab.data(), tab.data() + 2 * lw_);
    float *t2 = tab.data() + 2 * ((size_t)lw_ + lh);
    lw_map_trig(gw, gh, t2, t2 + 2 * gw);
    DevBuf<float> cache, dmaps, dtab;
    const uint32_t entries = (n + 1) * (n + 1);
    cache.alloc((size_t)entries * 3);
    dmaps.upload(maps, maps_len, b->stream);
    dtab.upload(tab.data(), tab.size(), b->stream);
    const float2 *ft = reinterpret_cast<const float2 *>(dtab.ptr);
    HIP_TRY(hipEventRecord(b->ev[0], b->stream));
    k_sky_cache<<<dim3((entries + kBlock - 1) / kBlock), dim3(kBlock), 0, b->stream>>>(*params, n, cache.ptr);
    HIP_TRY(hipGetLastError());
    k_blend_sky<<<dim3((unsigned)((lens_n + kBlock - 1) / kBlock)), dim3(kBlock), 0, b->stream>>>(
        dmaps.ptr, lw_, lh, ft, ft + lw_, cache.ptr, *params, weight);
    HIP_TRY(hipGetLastError());
    k_blend_sky<<<dim3((unsigned)((light_n + kBlock - 1) / kBlock)), dim3(kBlock), 0, b->stream>>>(
        dmaps.ptr + lens_n * 3, gw, gh, ft + lw_ + lh, ft + lw_ + lh + gw, cache.ptr, *params, weight);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(b->ev[1], b->stream));
    HIP_TRY(hipMemcpyAsync(maps, dmaps.ptr, maps_len * sizeof(float), hipMemcpyDeviceToHost, b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    float ms = 0.0f;
    HIP_TRY(hipEventElapsedTime(&ms, b->ev[0], b->ev[1]));
    lw_timing tm{};
    tm.trace_ms = ms;
    tm.launches = 3;
    tm.samples = lens_n + light_n;
    b->timing = tm;
  });
}

int lw_hip_game_sky_maps(lw_hip_backend *b, const lw_quick_sky *lens, const lw_quick_sky *light,
                         float sky_light_factor, const float sky_light_color[3], const uint32_t *atlas,
                         size_t atlas_len, uint32_t atlas_w, const uint32_t *stars, size_t stars_len,
                         float *out, size_t cap, lw_game_sky *info) {
  if (!b || !lens || !light || !sky_light_color || !atlas || !stars || !out || !info) return LW_EINVAL;
  return guarded(b, [&] {
    const uint32_t lw_ = LW_LENS_MAP_W, lh = LW_LENS_MAP_H, gw = LW_LIGHT_MAP_W, gh = LW_LIGHT_MAP_H;
    const size_t lens_n = (size_t)lw_ * lh, light_n = (size_t)gw * gh, total = (lens_n + light_n) * 3;
    if (cap < total) throw InvalidScene{"destination smaller than the two maps"};
    if (atlas_len == 0 || atlas_w == 0 || atlas_len > 0xFFFFFFFFull) throw InvalidScene{"empty atlas"};
    // the star table: offsets per bin that stay inside its entries
    const size_t nbins = (size_t)LW_STAR_BINS_AZ * LW_STAR_BINS_EL;
    if (stars_len < nbins + 1) throw InvalidScene{"star table shorter than its bins"};
    for (size_t k = 0; k < nbins; k++)
      if (stars[k] > stars[k + 1]) throw InvalidScene{"star table bins out of order"};
    if (nbins + 1 + (size_t)stars[nbins] * LW_STAR_WORDS > stars_len) throw InvalidScene{"star table entries past its end"};
    HIP_TRY(hipSetDevice(b->device));
    const double setup0 = b->pow_fix[POW_DECODE].ready ? 0.0 : -1.0;
    const auto &pf = ensure_pow_fix(b, POW_DECODE);
    std::vector<float> tab(2 * ((size_t)lw_ + lh + gw + gh));
    lw_map_trig(lw_, lh, tab.data(), tab.data() + 2 * lw_);
    float *t2 = tab.data() + 2 * ((size_t)lw_ + lh);
    lw_map_trig(gw, gh, t2, t2 + 2 * gw);
    lw::DiscSectors sec;
    lw_disc_sector_trig(sec.s0, sec.c0, sec.s1, sec.c1);
    DevBuf<float> dmaps, dtab;
    DevBuf<uint32_t> datlas, dstars;
    dmaps.alloc(total);
    dtab.upload(tab.data(), tab.size(), b->stream);
    datlas.upload(atlas, atlas_len, b->stream);
    dstars.upload(stars, stars_len, b->stream);
    const float2 *ft = reinterpret_cast<const float2 *>(dtab.ptr);
    float *light_map = dmaps.ptr + lens_n * 3;
    HIP_TRY(hipEventRecord(b->ev[0], b->stream));
    k_quick_sky_map<<<dim3((unsigned)((lens_n + kBlock - 1) / kBlock)), dim3(kBlock), 0, b->stream>>>(
        dmaps.ptr, lw_, lh, ft, ft + lw_, *lens, sec, datlas.ptr, atlas_w, (uint32_t)atlas_len, dstars.ptr,
        pf.fix.ptr, pf.n);
    HIP_TRY(hipGetLastError());
    k_quick_sky_map<<<dim3((unsigned)((light_n + kBlock - 1) / kBlock)), dim3(kBlock), 0, b->stream>>>(
        light_map, gw, gh, ft + lw_ + lh, ft + lw_ + lh + gw, *light, sec, datlas.ptr, atlas_w,
        (uint32_t)atlas_len, dstars.ptr, pf.fix.ptr, pf.n);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(b->ev[1], b->stream));
    // The light's cosine-weighted mean over the upper hemisphere, summed in f64 on the host in
    // the reference's order (sky.rs:922-931), then the scale and the veil (sky.rs:932-942).
    std::vector<float> game(light_n * 3);
    HIP_TRY(hipMemcpyAsync(game.data(), light_map, light_n * 3 * sizeof(float), hipMemcpyDeviceToHost, b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    double acc[3] = {0.0, 0.0, 0.0}, weight = 0.0;
    for (size_t k = 0; k < light_n; k++) {
      const float *e = t2 + 2 * gw + 2 * (k / gw);
      const float dy = e[1];
      if (dy > 0.0f) {
        const float r = 1.0f - dy * dy;
        const double w = (double)(dy * std::sqrt(r > 0.0f ? r : 0.0f));
        for (int c = 0; c < 3; c++) acc[c] += (double)game[k * 3 + c] * w;
        weight += w;
      }
    }
    float mean[3], target[3];
    for (int c = 0; c < 3; c++) mean[c] = (float)(acc[c] / weight);
    lw_lightmap_sky(sky_light_factor, sky_light_color, LW_DEFAULT_BRIGHTNESS, target);
    for (int c = 0; c < 3; c++) target[c] = lw::libm_powf(target[c], kPowExponent[POW_DECODE]);
    float scale = LW_MAX_SKY_LIGHT;
    for (int c = 0; c < 3; c++)
      if (mean[c] > 0.0f) scale = std::min(scale, target[c] / mean[c]);
    float veil[3];
    for (int c = 0; c < 3; c++) {
      const float v = target[c] / scale - mean[c];
      veil[c] = std::isnan(v) ? 0.0f : (v > 0.0f ? v : 0.0f);
    }
    HIP_TRY(hipEventRecord(b->ev[2], b->stream));
    k_sky_veil<<<dim3((unsigned)((light_n + kBlock - 1) / kBlock)), dim3(kBlock), 0, b->stream>>>(
        light_map, (uint32_t)light_n, make_float3(veil[0], veil[1], veil[2]), scale);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(b->ev[3], b->stream));
    HIP_TRY(hipMemcpyAsync(out, dmaps.ptr, total * sizeof(float), hipMemcpyDeviceToHost, b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    float ms_draw = 0.0f, ms_veil = 0.0f;
    HIP_TRY(hipEventElapsedTime(&ms_draw, b->ev[0], b->ev[1]));
    HIP_TRY(hipEventElapsedTime(&ms_veil, b->ev[2], b->ev[3]));
    std::memset(info, 0, sizeof(*info));
    info->sizes[0] = (float)lw_;
    info->sizes[1] = (float)lh;
    info->sizes[2] = (float)gw;
    info->sizes[3] = (float)gh;
    for (int c = 0; c < 3; c++) {
      info->light[c] = target[c];
      info->mean[c] = mean[c];
      info->veil[c] = veil[c];
    }
    info->scale = scale;
    info->kernel_ms = (double)ms_draw + (double)ms_veil;
    info->setup_ms = setup0 < 0.0 ? pf.ms : 0.0;
    lw_timing tm{};
    tm.trace_ms = info->kernel_ms;
    tm.launches = 3;
    tm.samples = lens_n + light_n;
    b->timing = tm;
  });
}

int lw_hip_encode(lw_hip_backend *b, const float *x, size_t n, float *out, double *setup_ms) {
  if (!b || (n && (!x || !out))) return LW_EINVAL;
  return guarded(b, [&] {
    HIP_TRY(hipSetDevice(b->device));
    if (n > 0xFFFFFFFFull) throw InvalidScene{"too many values"};
    const auto &pf = ensure_pow_fix(b, POW_ENCODE);
    if (setup_ms) *setup_ms = pf.ms;
    DevBuf<float> in, res;
    in.upload(x, n, b->stream);
    res.alloc(std::max<size_t>(n, 1));
    if (n) {
      k_encode<<<dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, b->stream>>>(
          in.ptr, (uint32_t)n, pf.fix.ptr, pf.n, res.ptr);
      HIP_TRY(hipGetLastError());
      HIP_TRY(hipMemcpyAsync(out, res.ptr, n * sizeof(float), hipMemcpyDeviceToHost, b->stream));
    }
    HIP_TRY(hipStreamSynchronize(b->stream));
  });
}

int lw_hip_casts(lw_hip_backend *b, const float *x, size_t n, int32_t *out_i, uint32_t *out_u) {
  if (!b || (n && (!x || !out_i || !out_u))) return LW_EINVAL;
  return guarded(b, [&] {
    HIP_TRY(hipSetDevice(b->device));
    if (n > 0xFFFFFFFFull) throw InvalidScene{"too many values"};
    if (n == 0) return;
    DevBuf<float> in;
    DevBuf<int32_t> ri;
    DevBuf<uint32_t> ru;
    in.upload(x, n, b->stream);
    ri.alloc(n);
    ru.alloc(n);
    k_casts<<<dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, b->stream>>>(in.ptr, (uint32_t)n,
                                                                                         ri.ptr, ru.ptr);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(out_i, ri.ptr, n * sizeof(int32_t), hipMemcpyDeviceToHost, b->stream));
    HIP_TRY(hipMemcpyAsync(out_u, ru.ptr, n * sizeof(uint32_t), hipMemcpyDeviceToHost, b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
  });
}

int lw_hip_downsample(lw_hip_backend *b, const uint32_t *px, uint32_t w, uint32_t h, uint32_t k,
                      uint32_t *out) {
  if (!b || !px || !out) return LW_EINVAL;
  return guarded(b, [&] {
    HIP_TRY(hipSetDevice(b->device));
    const size_t n = (size_t)w * h;
    if (k <= 1) {
      std::memcpy(out, px, n * sizeof(uint32_t));
      return;
    }
    const uint32_t ow = w / k, oh = h / k;
    const size_t m = (size_t)ow * oh;
    DevBuf<uint32_t> in, res;
    in.upload(px, n, b->stream);
    res.alloc(std::max<size_t>(m, 1));
    if (m) {
      k_downsample<<<dim3((unsigned)((m + kBlock - 1) / kBlock)), dim3(kBlock), 0, b->stream>>>(
          in.ptr, w, k, ow, oh, res.ptr);
      HIP_TRY(hipGetLastError());
      HIP_TRY(hipMemcpyAsync(out, res.ptr, m * sizeof(uint32_t), hipMemcpyDeviceToHost, b->stream));
    }
    HIP_TRY(hipStreamSynchronize(b->stream));
  });
}

int lw_hip_render_hq_rgba8(lw_hip_backend *b, const lw_hip_scene *scene, const lw_params *params,
                           uint32_t spp, float exposure, uint32_t op, uint32_t *out_rgba8) {
  if (!b || !scene || !params || !out_rgba8) return LW_EINVAL;
  return guarded(b, [&] {
    const ToneDev t = make_tone(exposure, op, op, 1.0f);
    render_frame(b, *scene, *params, spp, nullptr, nullptr, nullptr);
    const lw_timing render = b->timing;
    tonemap_device(b, b->out.ptr, (size_t)params->frame[0] * params->frame[1], t, false, out_rgba8);
    const double tone_ms = b->timing.trace_ms;
    b->timing = render;
    b->timing.accumulate_ms += tone_ms;
  });
}

int lw_hip_copy_probe(lw_hip_backend *b, size_t bytes, double *ms) {
  if (!b || !ms || bytes < 16) return LW_EINVAL;
  return guarded(b, [&] {
    HIP_TRY(hipSetDevice(b->device));
    const size_t n = bytes / 16;
    DevBuf<uint4> src, dst;
    src.alloc(n);
    dst.alloc(n);
    k_copy<<<dim3(2048), dim3(kBlock), 0, b->stream>>>(src.ptr, dst.ptr, n);  // warm
    HIP_TRY(hipEventRecord(b->ev[0], b->stream));
    k_copy<<<dim3(2048), dim3(kBlock), 0, b->stream>>>(src.ptr, dst.ptr, n);
    HIP_TRY(hipEventRecord(b->ev[1], b->stream));
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventSynchronize(b->ev[1]));
    float t = 0.0f;
    HIP_TRY(hipEventElapsedTime(&t, b->ev[0], b->ev[1]));
    *ms = t;
  });
}

int lw_hip_last_timing(const lw_hip_backend *b, lw_timing *out) {
  if (!b || !out) return LW_EINVAL;
  *out = b->timing;
  return LW_OK;
}

int lw_hip_trace(lw_hip_backend *b, const lw_scene_view *view, size_t n, const float *o,
                 const float *d, const float *tmin, const float *tmax, lw_hit *out) {
  if (!b || !view || (n && (!o || !d || !tmin || !tmax || !out))) return LW_EINVAL;
  return guarded(b, [&] {
    if (n > 0xFFFFFFFFull) throw InvalidScene{"too many rays"};
    if (n == 0) return;
    lw_hip_scene sc;
    upload_scene(b, *view, sc);
    DevBuf<float> d_o, d_d, d_tmin, d_tmax;
    DevBuf<lw_hit> d_out;
    d_o.upload(o, 3 * n, b->stream);
    d_d.upload(d, 3 * n, b->stream);
    d_tmin.upload(tmin, n, b->stream);
    d_tmax.upload(tmax, n, b->stream);
    d_out.alloc(n);
    k_trace<<<dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, b->stream>>>(
        *view->params, sc.dev(), (uint32_t)n, d_o.ptr, d_d.ptr, d_tmin.ptr, d_tmax.ptr, d_out.ptr);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(out, d_out.ptr, n * sizeof(lw_hit), hipMemcpyDeviceToHost, b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
  });
}

int lw_hip_render_samples(lw_hip_backend *b, const lw_scene_view *view, size_t n,
                          const uint32_t *px, const uint32_t *py, const uint32_t *sample,
                          float *out_rgb) {
  if (!b || !view || (n && (!px || !py || !sample || !out_rgb))) return LW_EINVAL;
  return guarded(b, [&] {
    if (n > 0xFFFFFFFFull) throw InvalidScene{"too many samples"};
    if (n == 0) return;
    lw_hip_scene sc;
    upload_scene(b, *view, sc);
    DevBuf<uint32_t> d_px, d_py, d_s;
    DevBuf<float> d_out;
    d_px.upload(px, n, b->stream);
    d_py.upload(py, n, b->stream);
    d_s.upload(sample, n, b->stream);
    d_out.alloc(3 * n);
    if (b->mode == MODE_WAVEFRONT) {
      const ListWork lw{d_px.ptr, d_py.ptr, d_s.ptr, d_out.ptr, (uint32_t)n};
      run_wavefront(b, sc, *view->params, lw, lw.total);
    } else if (b->mode == MODE_DIRECT) {
      k_render_samples<<<dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, b->stream>>>(
          *view->params, sc.dev(), (uint32_t)n, d_px.ptr, d_py.ptr, d_s.ptr, d_out.ptr);
    } else {
      const ListWork lw{d_px.ptr, d_py.ptr, d_s.ptr, d_out.ptr, (uint32_t)n};
      const uint32_t grid = std::min<uint32_t>(b->blocks_list, (uint32_t)((n + kBlock - 1) / kBlock));
      k_fill_u32<<<dim3(1), dim3(64), 0, b->stream>>>(b->counter.ptr, 0u);
      LW_WAVES_DISPATCH(b->waves, (k_paths<ListWork, W><<<dim3(grid), dim3(kBlock), 0, b->stream>>>(
                                      *view->params, sc.dev(), lw, b->counter.ptr)));
    }
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(out_rgb, d_out.ptr, 3 * n * sizeof(float), hipMemcpyDeviceToHost,
                           b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
  });
}

}  // extern "C"
