// The device scene builder: build_hq's layout (crates/lumenwright/src/scene/voxel.rs:305-491)
// and emitter_list (voxel.rs:542-607), made in HBM from the interned voxel volume.
//
// Kernels, in order:
//   k_classify    one workgroup per brick: whether any voxel is set, whether all equal voxel
//                 0, whether every voxel's model occludes, the first set voxel, the emitting
//                 voxels and the 4x4x4 sub-block occupancy (two 16-voxel rows per lane, read
//                 as 32 B)
//   k_decide      one lane per brick: EMPTY, buried, UNIFORM or stored (voxel.rs:307-341)
//   scans         stored bricks -> brick slots; emitting voxels -> emitter indices
//   k_sites       one workgroup per brick with emitters: its emitters in voxel order
//                 (voxel.rs:315-324, 553-569) and the bounds of their grid cells
//   k_cell_keys   one lane per emitter: its 27 (cell, emitter) keys (voxel.rs:585-595);
//                 a radix sort then orders them by cell, emitters ascending in a cell
//   k_cells       one lane per cell: (start, count) of its list (voxel.rs:597-606)
//   k_ov_*        the entity overlay (voxel.rs:378-425): each quad's cells as (cell, quad)
//                 keys, sorted, then the overlay bricks, cells and lists
//   k_store       one workgroup per brick: its top-level entry and, for a stored brick, its
//                 2048 words and occupancy
// All index arithmetic on the host side is checked against the 2^32-word limits of the
// layout before the buffers are allocated.
#include "lw_build.hpp"

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <algorithm>
#include <chrono>
#include <climits>
#include <vector>

namespace lwb {
namespace {

constexpr uint32_t kB = 256;
constexpr uint32_t kEmpty = 0xFFFFFFFFu, kUniform = 0x80000000u;
constexpr int32_t kEmitterCell = 10;  // voxel.rs:532

#define LWB_TRY(expr)                                                                        \
  do {                                                                                       \
    hipError_t e_ = (expr);                                                                  \
    if (e_ != hipSuccess) throw Error{false, e_, std::string(#expr) + ": " + hipGetErrorString(e_)}; \
  } while (0)

[[noreturn]] void invalid(const std::string &m) { throw Error{true, hipSuccess, m}; }

template <class T>
struct Buf {
  T *p = nullptr;
  size_t n = 0;
  void alloc(size_t count) {
    free();
    if (count) {
      LWB_TRY(hipMalloc((void **)&p, count * sizeof(T)));
      n = count;
    }
  }
  T *release() {
    T *r = p;
    p = nullptr;
    n = 0;
    return r;
  }
  void free() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  ~Buf() { free(); }
};

enum : uint32_t { B_SET = 1u, B_MIXED = 2u, B_OCCLUDES = 4u, B_BAD_ID = 8u };
enum : uint32_t { K_EMPTY = 0u, K_BURIED = 1u, K_UNIFORM = 2u, K_STORED = 3u };

struct BrickSummary {
  uint32_t bits;
  uint32_t first;  // (local << 16) | id of the first set voxel, 0xFFFFFFFF if none
  uint32_t id0;    // voxel 0's id
  uint32_t emit;   // emitting voxels
};

uint32_t bit_width(uint64_t v) {
  uint32_t b = 0;
  while (v) {
    b++;
    v >>= 1;
  }
  return b;
}

// ---- kernels ----

template <class T>
__global__ void __launch_bounds__(kB) k_fill(T *__restrict__ p, size_t n, T v) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    p[i] = v;
}

__global__ void __launch_bounds__(kB) k_classify(const uint16_t *__restrict__ vox,
                                                 const uint8_t *__restrict__ occ_in,
                                                 const uint8_t *__restrict__ mflags,
                                                 uint32_t models_len, BrickSummary *__restrict__ sum,
                                                 uint2 *__restrict__ occm) {
  __shared__ uint32_t s_bits, s_first, s_emit, s_notocc, s_m[2];
  const uint32_t b = blockIdx.x, t = threadIdx.x;
  if (t == 0) {
    s_bits = 0u;
    s_first = 0xFFFFFFFFu;
    s_emit = 0u;
    s_notocc = 0u;
    s_m[0] = 0u;
    s_m[1] = 0u;
  }
  __syncthreads();
  // Lane t holds the 16 voxels of row (ly, lz) = (t / 16, t % 16), lx = 0..15.
  const uint16_t *brick = vox + (size_t)b * 4096;
  const uint4 *row = reinterpret_cast<const uint4 *>(brick + t * 16);
  const uint4 r0 = row[0], r1 = row[1];
  const uint32_t w[8] = {r0.x, r0.y, r0.z, r0.w, r1.x, r1.y, r1.z, r1.w};
  const uint32_t id0 = brick[0];
  uint32_t bits = 0u, first = 0xFFFFFFFFu, emit = 0u, quarters = 0u;
  bool notocc = false;
#pragma unroll
  for (int k = 0; k < 16; k++) {
    const uint32_t id = (w[k >> 1] >> ((k & 1) * 16)) & 0xFFFFu;
    if (id != id0) bits |= B_MIXED;
    if (id >= models_len) {
      bits |= B_BAD_ID;
      continue;
    }
    const uint32_t f = mflags[id];
    if (!(f & LW_MODEL_OCCLUDES)) notocc = true;
    if (f & LW_MODEL_EMITS) emit++;
    if (id != 0u) {
      bits |= B_SET;
      quarters |= 1u << (k >> 2);
      if (first == 0xFFFFFFFFu) first = ((t * 16u + (uint32_t)k) << 16) | id;
    }
  }
  // Sub-block ((ly / 4) * 4 + lz / 4) * 4 + lx / 4 (the occupancy's bit order).
  const uint32_t sb = ((t >> 6) * 4u + ((t & 15u) >> 2)) * 4u;
  if (bits) atomicOr(&s_bits, bits);
  if (first != 0xFFFFFFFFu) atomicMin(&s_first, first);
  if (emit) atomicAdd(&s_emit, emit);
  if (quarters) atomicOr(&s_m[sb >> 5], quarters << (sb & 31u));
  if (notocc) atomicOr(&s_notocc, 1u);
  __syncthreads();
  if (t == 0) {
    uint32_t bb = s_bits;
    const bool occl = occ_in ? occ_in[b] != 0 : s_notocc == 0u;
    if (occl) bb |= B_OCCLUDES;
    sum[b] = BrickSummary{bb, s_first, id0, s_emit};
    occm[b] = make_uint2(s_m[0], s_m[1]);
  }
}

// voxel.rs:307-341. counts: [0] uniform, [1] buried, [2] bricks with an unknown id.
__global__ void __launch_bounds__(kB) k_decide(const BrickSummary *__restrict__ sum, uint32_t dx,
                                               uint32_t dy, uint32_t dz, uint32_t nb,
                                               uint32_t *__restrict__ kind, uint32_t *__restrict__ stored,
                                               uint32_t *__restrict__ ecount,
                                               unsigned long long *__restrict__ counts) {
  for (uint32_t b = blockIdx.x * blockDim.x + threadIdx.x; b < nb; b += gridDim.x * blockDim.x) {
    const BrickSummary s = sum[b];
    if (s.bits & B_BAD_ID) atomicAdd(&counts[2], 1ull);
    uint32_t k = K_EMPTY;
    if (s.bits & B_SET) {
      const int32_t bx = (int32_t)(b % dx), by = (int32_t)(b / (dx * dz)), bz = (int32_t)((b / dx) % dz);
      auto occludes = [&](int32_t x, int32_t y, int32_t z) {
        if (x < 0 || y < 0 || z < 0 || x >= (int32_t)dx || y >= (int32_t)dy || z >= (int32_t)dz) return false;
        return (sum[((uint32_t)y * dz + (uint32_t)z) * dx + (uint32_t)x].bits & B_OCCLUDES) != 0u;
      };
      const bool buried = (s.bits & B_OCCLUDES) && occludes(bx + 1, by, bz) && occludes(bx - 1, by, bz) &&
                          occludes(bx, by + 1, bz) && occludes(bx, by - 1, bz) && occludes(bx, by, bz + 1) &&
                          occludes(bx, by, bz - 1);
      k = buried ? K_BURIED : ((s.bits & B_MIXED) ? K_STORED : K_UNIFORM);
      if (k == K_BURIED) atomicAdd(&counts[1], 1ull);
      if (k == K_UNIFORM) atomicAdd(&counts[0], 1ull);
    }
    kind[b] = k;
    stored[b] = k == K_STORED ? 1u : 0u;
    ecount[b] = (s.bits & B_SET) ? s.emit : 0u;  // an empty section has no instances
  }
}

// The emitters of one brick in voxel order (voxel.rs:315-324) as GpuEmitter records
// (voxel.rs:553-569), and the bounds of their cells. bounds: lo x, y, z then hi x, y, z.
__global__ void __launch_bounds__(kB) k_sites(const uint16_t *__restrict__ vox,
                                              const uint32_t *__restrict__ ecount,
                                              const unsigned long long *__restrict__ eoff,
                                              const uint8_t *__restrict__ mflags,
                                              const lw_model *__restrict__ models, uint32_t dx,
                                              uint32_t dz, lw_emitter *__restrict__ out,
                                              int32_t *__restrict__ bounds) {
  const uint32_t b = blockIdx.x, t = threadIdx.x;
  if (ecount[b] == 0u) return;  // the whole workgroup leaves
  __shared__ uint32_t s_wave[kB / 64];
  __shared__ int32_t s_lo[3], s_hi[3];
  if (t < 3) {
    s_lo[t] = INT_MAX;
    s_hi[t] = INT_MIN;
  }
  const uint4 *row = reinterpret_cast<const uint4 *>(vox + (size_t)b * 4096 + t * 16);
  const uint4 r0 = row[0], r1 = row[1];
  const uint32_t w[8] = {r0.x, r0.y, r0.z, r0.w, r1.x, r1.y, r1.z, r1.w};
  uint32_t mask = 0u;
#pragma unroll
  for (int k = 0; k < 16; k++) {
    const uint32_t id = (w[k >> 1] >> ((k & 1) * 16)) & 0xFFFFu;
    if (mflags[id] & LW_MODEL_EMITS) mask |= 1u << k;  // ids checked by k_classify
  }
  const uint32_t c = (uint32_t)__popc(mask);
  const uint32_t lane = t & 63u, wave = t >> 6;
  uint32_t incl = c;
#pragma unroll
  for (uint32_t d = 1; d < 64; d <<= 1) {
    const uint32_t v = __shfl_up(incl, d);
    if (lane >= d) incl += v;
  }
  if (lane == 63u) s_wave[wave] = incl;
  __syncthreads();
  uint32_t before = 0u;
  for (uint32_t i = 0; i < wave; i++) before += s_wave[i];
  unsigned long long idx = eoff[b] + before + incl - c;
  const int32_t bx = (int32_t)(b % dx), by = (int32_t)(b / (dx * dz)), bz = (int32_t)((b / dx) % dz);
  const int32_t y = by * 16 + (int32_t)(t >> 4), z = bz * 16 + (int32_t)(t & 15u);
  int32_t lo[3] = {INT_MAX, INT_MAX, INT_MAX}, hi[3] = {INT_MIN, INT_MIN, INT_MIN};
  for (uint32_t m = mask; m; m &= m - 1u) {
    const int k = __ffs(m) - 1;
    const uint32_t id = (w[k >> 1] >> ((k & 1) * 16)) & 0xFFFFu;
    const int32_t x = bx * 16 + k;
    const lw_model md = models[id];
    lw_emitter e;
    e.pos[0] = (float)x;
    e.pos[1] = (float)y;
    e.pos[2] = (float)z;
    e.pos[3] = __uint_as_float(md.quad_offset);
    e.faces[0] = md.quad_count;
    e.faces[1] = 0u;
    e.faces[2] = 0u;
    e.faces[3] = 0u;
    out[idx++] = e;
    const int32_t g[3] = {x / kEmitterCell, y / kEmitterCell, z / kEmitterCell};  // x, y, z >= 0
    for (int a = 0; a < 3; a++) {
      lo[a] = min(lo[a], g[a]);
      hi[a] = max(hi[a], g[a]);
    }
  }
  if (c) {
    for (int a = 0; a < 3; a++) {
      atomicMin(&s_lo[a], lo[a]);
      atomicMax(&s_hi[a], hi[a]);
    }
  }
  __syncthreads();
  if (t < 3) {
    atomicMin(&bounds[t], s_lo[t]);
    atomicMax(&bounds[3 + t], s_hi[t]);
  }
}

struct CellFrame {
  int32_t offset[3];
  uint32_t size[3];
};

// Each emitter in its own cell and the 26 around it (voxel.rs:585-595): keys cell << 32 | i.
__global__ void __launch_bounds__(kB) k_cell_keys(const lw_emitter *__restrict__ em, uint32_t n,
                                                  CellFrame f, unsigned long long *__restrict__ keys,
                                                  uint32_t *__restrict__ counts) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    int32_t g[3];
    for (int a = 0; a < 3; a++) g[a] = (int32_t)em[i].pos[a] / kEmitterCell;
    uint32_t k = 0;
    for (int dy = -1; dy <= 1; dy++)
      for (int dz = -1; dz <= 1; dz++)
        for (int dx = -1; dx <= 1; dx++) {
          const uint32_t l0 = (uint32_t)(g[0] + dx - f.offset[0]), l1 = (uint32_t)(g[1] + dy - f.offset[1]),
                         l2 = (uint32_t)(g[2] + dz - f.offset[2]);
          const uint32_t cell = (l1 * f.size[2] + l2) * f.size[0] + l0;
          keys[(size_t)i * 27 + k++] = ((unsigned long long)cell << 32) | i;
          atomicAdd(&counts[cell], 1u);
        }
  }
}

__global__ void __launch_bounds__(kB) k_cells(const uint32_t *__restrict__ counts,
                                              const uint32_t *__restrict__ starts, uint32_t ncells,
                                              uint32_t *__restrict__ egrid, uint32_t *__restrict__ longest) {
  uint32_t most = 0u;
  for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < ncells; c += gridDim.x * blockDim.x) {
    egrid[LW_EMITTER_GRID_HEADER + 2 * (size_t)c] = starts[c];
    egrid[LW_EMITTER_GRID_HEADER + 2 * (size_t)c + 1] = counts[c];
    most = max(most, counts[c]);
  }
  if (most) atomicMax(longest, most);
}

// The emitter grid's header (voxel.rs:597-605).
__global__ void k_cell_header(uint32_t *__restrict__ egrid, CellFrame f) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  egrid[0] = (uint32_t)kEmitterCell;
  for (int a = 0; a < 3; a++) {
    egrid[1 + a] = f.size[a];
    egrid[4 + a] = (uint32_t)f.offset[a];
  }
  egrid[7] = 0u;
}

__global__ void __launch_bounds__(kB) k_cell_list(const unsigned long long *__restrict__ keys, size_t n,
                                                  uint32_t *__restrict__ list) {
  for (size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += (size_t)gridDim.x * blockDim.x)
    list[j] = (uint32_t)keys[j];
}

// Rust `as i32` of a float: saturating, NaN -> 0.
__device__ __forceinline__ int32_t as_i32(float x) {
  if (isnan(x)) return 0;
  if (x >= 2147483648.0f) return INT_MAX;
  if (x <= -2147483648.0f) return INT_MIN;
  return (int32_t)x;
}

struct OvFrame {
  int32_t size[3];  // the grid in blocks
  uint32_t dx, dz;
  uint32_t qbits;   // bits of the quad index in a key
};

// An overlay quad's voxel box (voxel.rs:390-405): its corners' bounds grown by 1e-3, floored,
// clamped to the grid. cells: the box's voxel count (0 when empty).
__global__ void __launch_bounds__(kB) k_ov_boxes(const lw_quad *__restrict__ q, uint32_t n, OvFrame f,
                                                 int4 *__restrict__ box_lo, int4 *__restrict__ box_hi,
                                                 unsigned long long *__restrict__ cells) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const lw_quad Q = q[i];
    int32_t l[3], h[3];
    unsigned long long count = 1ull;
    for (int a = 0; a < 3; a++) {
      const float c0 = Q.p0[a], c1 = Q.p0[a] + Q.ev[a], c2 = (Q.p0[a] + Q.eu[a]) + Q.ev[a],
                  c3 = Q.p0[a] + Q.eu[a];
      float lo = fminf(fminf(fminf(fminf(3.40282347e38f, c0), c1), c2), c3);
      float hi = fmaxf(fmaxf(fmaxf(fmaxf(-3.40282347e38f, c0), c1), c2), c3);
      lo = lo - 1e-3f;
      hi = hi + 1e-3f;
      l[a] = max(as_i32(floorf(lo)), 0);
      h[a] = min(as_i32(floorf(hi)), f.size[a] - 1);
      count = h[a] >= l[a] ? count * (unsigned long long)(h[a] - l[a] + 1) : 0ull;
    }
    box_lo[i] = make_int4(l[0], l[1], l[2], 0);
    box_hi[i] = make_int4(h[0], h[1], h[2], 0);
    cells[i] = count;
  }
}

// The (cell, quad) keys of one quad, cell = brick * 4096 + local (voxel.rs:406-423).
__global__ void __launch_bounds__(kB) k_ov_keys(const int4 *__restrict__ box_lo, const int4 *__restrict__ box_hi,
                                                const unsigned long long *__restrict__ off, uint32_t n,
                                                OvFrame f, unsigned long long *__restrict__ keys) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int4 l = box_lo[i], h = box_hi[i];
    unsigned long long j = off[i];
    for (int32_t y = l.y; y <= h.y; y++)
      for (int32_t z = l.z; z <= h.z; z++)
        for (int32_t x = l.x; x <= h.x; x++) {
          const unsigned long long b = ((unsigned long long)(y >> 4) * f.dz + (unsigned long long)(z >> 4)) * f.dx +
                                       (unsigned long long)(x >> 4);
          const unsigned long long local = (unsigned long long)(((y & 15) * 16 + (z & 15)) * 16 + (x & 15));
          keys[j++] = ((b * 4096ull + local) << f.qbits) | i;
        }
  }
}

__global__ void __launch_bounds__(kB) k_ov_marks(const unsigned long long *__restrict__ keys, size_t n,
                                                 uint32_t qbits, uint32_t *__restrict__ first) {
  for (size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += (size_t)gridDim.x * blockDim.x)
    first[j] = (j == 0 || (keys[j] >> (qbits + 12)) != (keys[j - 1] >> (qbits + 12))) ? 1u : 0u;
}

// Overlay bricks ascending (voxel.rs:413-416); each cell's run of the sorted keys is its list.
__global__ void __launch_bounds__(kB) k_ov_fill(const unsigned long long *__restrict__ keys, size_t n,
                                                uint32_t qbits, const uint32_t *__restrict__ first,
                                                const uint32_t *__restrict__ rank, uint32_t entity_base,
                                                uint32_t *__restrict__ ov_top, uint32_t *__restrict__ ov_cells,
                                                uint32_t *__restrict__ ov_quads) {
  for (size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += (size_t)gridDim.x * blockDim.x) {
    const unsigned long long cell = keys[j] >> qbits;
    const uint32_t r = rank[j] + first[j] - 1u;
    if (first[j]) ov_top[cell >> 12] = r;
    ov_quads[j] = entity_base + (uint32_t)(keys[j] & ((1ull << qbits) - 1ull));
    if (j == 0 || (keys[j - 1] >> qbits) != cell)
      ov_cells[(size_t)r * LW_OV_CELL_WORDS + (cell & 4095ull) * 2] = (uint32_t)j;
  }
}

__global__ void __launch_bounds__(kB) k_ov_counts(const unsigned long long *__restrict__ keys, size_t n,
                                                  uint32_t qbits, const uint32_t *__restrict__ first,
                                                  const uint32_t *__restrict__ rank,
                                                  uint32_t *__restrict__ ov_cells) {
  for (size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += (size_t)gridDim.x * blockDim.x) {
    const unsigned long long cell = keys[j] >> qbits;
    if (j + 1 < n && (keys[j + 1] >> qbits) == cell) continue;
    const uint32_t r = rank[j] + first[j] - 1u;
    uint32_t *slot = ov_cells + (size_t)r * LW_OV_CELL_WORDS + (cell & 4095ull) * 2;
    slot[1] = (uint32_t)j + 1u - slot[0];
  }
}

// voxel.rs:325-341: the top-level entry; a stored brick's 2048 words are its 4096 ids in
// pairs (low half first), i.e. the ids' own little-endian bytes.
__global__ void __launch_bounds__(kB) k_store(const uint16_t *__restrict__ vox,
                                              const BrickSummary *__restrict__ sum,
                                              const uint32_t *__restrict__ kind,
                                              const uint32_t *__restrict__ slot,
                                              const uint2 *__restrict__ occm, uint32_t *__restrict__ grid,
                                              size_t off_bricks, uint2 *__restrict__ occupancy) {
  const uint32_t b = blockIdx.x, t = threadIdx.x;
  const uint32_t k = kind[b];
  if (t == 0) {
    const BrickSummary s = sum[b];
    uint32_t e = kEmpty;
    if (k == K_BURIED) e = kUniform | (s.first & 0xFFFFu);
    if (k == K_UNIFORM) e = kUniform | s.id0;
    if (k == K_STORED) e = slot[b];
    grid[b] = e;
    if (k == K_STORED) occupancy[slot[b]] = occm[b];
  }
  if (k != K_STORED) return;
  // The brick words start at off_bricks (the brick count), so only 4-byte alignment holds.
  const uint4 *src = reinterpret_cast<const uint4 *>(vox + (size_t)b * 4096) + t * 2;
  const uint4 r0 = src[0], r1 = src[1];
  uint32_t *dst = grid + off_bricks + (size_t)slot[b] * LW_BRICK_WORDS + t * 8;
  dst[0] = r0.x;
  dst[1] = r0.y;
  dst[2] = r0.z;
  dst[3] = r0.w;
  dst[4] = r1.x;
  dst[5] = r1.y;
  dst[6] = r1.z;
  dst[7] = r1.w;
}

unsigned grid_for(size_t n) { return (unsigned)std::max<size_t>(1, std::min<size_t>((n + kB - 1) / kB, 1u << 20)); }

template <class T>
void fill(T *p, size_t n, T v, hipStream_t s) {
  if (!n) return;
  k_fill<T><<<dim3(grid_for(n)), dim3(kB), 0, s>>>(p, n, v);
  LWB_TRY(hipGetLastError());
}

template <class In, class Out, class Acc>
void exclusive_scan(const In *in, Out *out, size_t n, Acc init, hipStream_t s) {
  size_t bytes = 0;
  LWB_TRY(rocprim::exclusive_scan(nullptr, bytes, in, out, init, n, rocprim::plus<Acc>(), s));
  Buf<char> tmp;
  tmp.alloc(std::max<size_t>(bytes, 1));
  LWB_TRY(rocprim::exclusive_scan(tmp.p, bytes, in, out, init, n, rocprim::plus<Acc>(), s));
  LWB_TRY(hipStreamSynchronize(s));  // tmp is freed on return
}

void sort_keys(const unsigned long long *in, unsigned long long *out, size_t n, uint32_t end_bit,
               hipStream_t s) {
  size_t bytes = 0;
  LWB_TRY(rocprim::radix_sort_keys(nullptr, bytes, in, out, n, 0, end_bit, s));
  Buf<char> tmp;
  tmp.alloc(std::max<size_t>(bytes, 1));
  LWB_TRY(rocprim::radix_sort_keys(tmp.p, bytes, in, out, n, 0, end_bit, s));
  LWB_TRY(hipStreamSynchronize(s));
}

template <class T>
T read_one(const T *p, hipStream_t s) {
  T v;
  LWB_TRY(hipMemcpyAsync(&v, p, sizeof(T), hipMemcpyDeviceToHost, s));
  LWB_TRY(hipStreamSynchronize(s));
  return v;
}

}  // namespace

void build(hipStream_t s, const Input &in, Output &out) {
  out = Output{};
  const uint64_t dx = in.dims[0], dy = in.dims[1], dz = in.dims[2];
  const uint64_t nb = dx * dy * dz;
  if (nb == 0 || nb >= (1ull << 24)) invalid("brick count out of range (1 .. 2^24 - 1)");

  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  LWB_TRY(hipEventCreate(&ev0));
  struct Events {
    hipEvent_t *a, *b;
    ~Events() {
      if (*a) (void)hipEventDestroy(*a);
      if (*b) (void)hipEventDestroy(*b);
    }
  } events{&ev0, &ev1};
  LWB_TRY(hipEventCreate(&ev1));
  LWB_TRY(hipEventRecord(ev0, s));

  // Bricks.
  Buf<BrickSummary> sum;
  Buf<uint2> occm;
  Buf<uint32_t> kind, stored, slot, ecount;
  Buf<unsigned long long> counts, eoff;
  sum.alloc(nb);
  occm.alloc(nb);
  kind.alloc(nb);
  stored.alloc(nb);
  slot.alloc(nb);
  ecount.alloc(nb);
  eoff.alloc(nb);
  counts.alloc(3);
  fill(counts.p, 3, 0ull, s);
  k_classify<<<dim3((unsigned)nb), dim3(kB), 0, s>>>(in.voxels, in.brick_occludes, in.model_flags,
                                                     in.models_len, sum.p, occm.p);
  LWB_TRY(hipGetLastError());
  k_decide<<<dim3(grid_for(nb)), dim3(kB), 0, s>>>(sum.p, (uint32_t)dx, (uint32_t)dy, (uint32_t)dz,
                                                   (uint32_t)nb, kind.p, stored.p, ecount.p, counts.p);
  LWB_TRY(hipGetLastError());
  exclusive_scan(stored.p, slot.p, nb, 0u, s);
  exclusive_scan(ecount.p, eoff.p, nb, 0ull, s);
  unsigned long long hc[3];
  LWB_TRY(hipMemcpyAsync(hc, counts.p, sizeof(hc), hipMemcpyDeviceToHost, s));
  const uint64_t n_stored = (uint64_t)read_one(slot.p + nb - 1, s) + read_one(stored.p + nb - 1, s);
  const uint64_t n_emit = (uint64_t)read_one(eoff.p + nb - 1, s) + read_one(ecount.p + nb - 1, s);
  if (hc[2]) invalid("a voxel names a model past models_len");
  out.uniform_bricks = hc[0];
  out.buried_bricks = hc[1];
  out.bricks = n_stored;
  out.emitters_count = n_emit;

  // Emitters and the emitter grid (voxel.rs:542-607).
  Buf<lw_emitter> emitters;
  Buf<uint32_t> egrid;
  uint32_t longest = 0;
  if (n_emit == 0) {
    emitters.alloc(1);
    LWB_TRY(hipMemsetAsync(emitters.p, 0, sizeof(lw_emitter), s));
    egrid.alloc(LW_EMITTER_GRID_HEADER);
    fill(egrid.p, LW_EMITTER_GRID_HEADER, 0u, s);
    fill(egrid.p, 1, (uint32_t)kEmitterCell, s);
  } else {
    if (27 * n_emit >= (1ull << 32)) invalid("too many emitters for the emitter grid");
    emitters.alloc(n_emit);
    Buf<int32_t> bounds;
    bounds.alloc(6);
    fill(bounds.p, 3, (int32_t)INT_MAX, s);
    fill(bounds.p + 3, 3, (int32_t)INT_MIN, s);
    k_sites<<<dim3((unsigned)nb), dim3(kB), 0, s>>>(in.voxels, ecount.p, eoff.p, in.model_flags, in.models,
                                                    (uint32_t)dx, (uint32_t)dz, emitters.p, bounds.p);
    LWB_TRY(hipGetLastError());
    int32_t hb[6];
    LWB_TRY(hipMemcpyAsync(hb, bounds.p, sizeof(hb), hipMemcpyDeviceToHost, s));
    LWB_TRY(hipStreamSynchronize(s));
    CellFrame f;
    uint64_t ncells = 1;
    for (int a = 0; a < 3; a++) {
      f.offset[a] = hb[a] - 1;
      f.size[a] = (uint32_t)(hb[3 + a] - hb[a] + 3);
      ncells *= f.size[a];
    }
    const uint64_t n_keys = 27 * n_emit;
    const uint64_t egrid_len = LW_EMITTER_GRID_HEADER + 2 * ncells + n_keys;
    if (egrid_len > 0xFFFFFFFFull) invalid("emitter grid past 2^32 words");
    egrid.alloc(egrid_len);
    Buf<uint32_t> ccount, cstart, dlong;
    Buf<unsigned long long> keys, sorted;
    ccount.alloc(ncells);
    cstart.alloc(ncells);
    dlong.alloc(1);
    keys.alloc(n_keys);
    sorted.alloc(n_keys);
    fill(ccount.p, ncells, 0u, s);
    fill(dlong.p, 1, 0u, s);
    k_cell_keys<<<dim3(grid_for(n_emit)), dim3(kB), 0, s>>>(emitters.p, (uint32_t)n_emit, f, keys.p, ccount.p);
    LWB_TRY(hipGetLastError());
    sort_keys(keys.p, sorted.p, n_keys, 32 + std::max(1u, bit_width(ncells - 1)), s);
    exclusive_scan(ccount.p, cstart.p, ncells, 0u, s);
    k_cell_header<<<dim3(1), dim3(64), 0, s>>>(egrid.p, f);
    LWB_TRY(hipGetLastError());
    k_cells<<<dim3(grid_for(ncells)), dim3(kB), 0, s>>>(ccount.p, cstart.p, (uint32_t)ncells, egrid.p, dlong.p);
    LWB_TRY(hipGetLastError());
    k_cell_list<<<dim3(grid_for(n_keys)), dim3(kB), 0, s>>>(sorted.p, n_keys,
                                                           egrid.p + LW_EMITTER_GRID_HEADER + 2 * ncells);
    LWB_TRY(hipGetLastError());
    longest = read_one(dlong.p, s);
  }

  // The entity overlay (voxel.rs:378-425).
  uint64_t n_ov_keys = 0, n_ov_bricks = 0;
  uint32_t qbits = 1;
  Buf<unsigned long long> ov_sorted;
  Buf<uint32_t> ov_first, ov_rank;
  if (in.n_entity > 0) {
    qbits = std::max(1u, bit_width(in.n_entity - 1));
    if (bit_width(nb * 4096 - 1) + qbits > 64) invalid("too many overlay quads for the key");
    OvFrame f{{(int32_t)(dx * 16), (int32_t)(dy * 16), (int32_t)(dz * 16)}, (uint32_t)dx, (uint32_t)dz, qbits};
    Buf<int4> lo, hi;
    Buf<unsigned long long> cells, coff, keys;
    lo.alloc(in.n_entity);
    hi.alloc(in.n_entity);
    cells.alloc(in.n_entity);
    coff.alloc(in.n_entity);
    k_ov_boxes<<<dim3(grid_for(in.n_entity)), dim3(kB), 0, s>>>(in.entity_quads, in.n_entity, f, lo.p, hi.p,
                                                                cells.p);
    LWB_TRY(hipGetLastError());
    exclusive_scan(cells.p, coff.p, in.n_entity, 0ull, s);
    n_ov_keys = read_one(coff.p + in.n_entity - 1, s) + read_one(cells.p + in.n_entity - 1, s);
    if (n_ov_keys >= (1ull << 31)) invalid("overlay quads cover more than 2^31 voxel cells");
    if (n_ov_keys > 0) {
      keys.alloc(n_ov_keys);
      ov_sorted.alloc(n_ov_keys);
      ov_first.alloc(n_ov_keys);
      ov_rank.alloc(n_ov_keys);
      k_ov_keys<<<dim3(grid_for(in.n_entity)), dim3(kB), 0, s>>>(lo.p, hi.p, coff.p, in.n_entity, f, keys.p);
      LWB_TRY(hipGetLastError());
      sort_keys(keys.p, ov_sorted.p, n_ov_keys, bit_width(nb * 4096 - 1) + qbits, s);
      k_ov_marks<<<dim3(grid_for(n_ov_keys)), dim3(kB), 0, s>>>(ov_sorted.p, n_ov_keys, qbits, ov_first.p);
      LWB_TRY(hipGetLastError());
      exclusive_scan(ov_first.p, ov_rank.p, n_ov_keys, 0u, s);
      n_ov_bricks = (uint64_t)read_one(ov_rank.p + n_ov_keys - 1, s) + read_one(ov_first.p + n_ov_keys - 1, s);
    }
  }
  out.overlay_bricks = n_ov_bricks;

  // The grid buffer (voxel.rs:427-440).
  const uint64_t off_bricks = nb, off_ov_top = off_bricks + n_stored * LW_BRICK_WORDS,
                 off_ov_cells = off_ov_top + nb, off_ov_quads = off_ov_cells + n_ov_bricks * LW_OV_CELL_WORDS,
                 grid_len = off_ov_quads + n_ov_keys + 1;
  if (grid_len > 0xFFFFFFFFull) invalid("grid buffer past 2^32 words");
  Buf<uint32_t> grid;
  Buf<uint2> occupancy;
  grid.alloc(grid_len);
  occupancy.alloc(std::max<uint64_t>(n_stored, 1));
  if (n_stored == 0) fill(occupancy.p, 1, make_uint2(0u, 0u), s);
  k_store<<<dim3((unsigned)nb), dim3(kB), 0, s>>>(in.voxels, sum.p, kind.p, slot.p, occm.p, grid.p, off_bricks,
                                                  occupancy.p);
  LWB_TRY(hipGetLastError());
  fill(grid.p + off_ov_top, nb, kEmpty, s);
  fill(grid.p + off_ov_cells, n_ov_bricks * LW_OV_CELL_WORDS, 0u, s);
  fill(grid.p + grid_len - 1, 1, 0u, s);
  if (n_ov_keys > 0) {
    k_ov_fill<<<dim3(grid_for(n_ov_keys)), dim3(kB), 0, s>>>(ov_sorted.p, n_ov_keys, qbits, ov_first.p, ov_rank.p,
                                                             in.entity_base, grid.p + off_ov_top,
                                                             grid.p + off_ov_cells, grid.p + off_ov_quads);
    LWB_TRY(hipGetLastError());
    k_ov_counts<<<dim3(grid_for(n_ov_keys)), dim3(kB), 0, s>>>(ov_sorted.p, n_ov_keys, qbits, ov_first.p,
                                                               ov_rank.p, grid.p + off_ov_cells);
    LWB_TRY(hipGetLastError());
  }
  LWB_TRY(hipEventRecord(ev1, s));
  LWB_TRY(hipStreamSynchronize(s));
  float ms = 0.0f;
  LWB_TRY(hipEventElapsedTime(&ms, ev0, ev1));

  out.offsets[0] = 0;
  out.offsets[1] = (uint32_t)off_bricks;
  out.offsets[2] = (uint32_t)off_ov_top;
  out.offsets[3] = (uint32_t)off_ov_cells;
  out.offsets[4] = (uint32_t)off_ov_quads;
  out.max_cell_list = longest;
  out.kernel_ms = ms;
  out.grid_len = grid_len;
  out.grid = grid.release();
  out.emitters_len = emitters.n;
  out.emitters = emitters.release();
  out.egrid_len = egrid.n;
  out.egrid = egrid.release();
  out.occupancy_len = occupancy.n;
  out.occupancy = occupancy.release();
}

}  // namespace lwb
