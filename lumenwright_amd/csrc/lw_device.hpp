// The HQ path tracer on gfx950: device functions for one path sample.
//
// The computation is the reference kernel crate's (crates/lumenwright-kernel/src/
// trace.rs and integrator.rs), operation for operation, so a ray meets the same quad at
// the same t bits and a sample returns the same radiance. The data access is the GPU's
// own:
//   * Params arrive as a kernel argument (scalar loads, uniform across the wave);
//   * the top-level brick entries are re-read only when the DDA enters a new brick;
//   * quads carry a precomputed axis code (QuadInfo): a quad whose edges are unit axis
//     vectors -- every face of a full cube -- is intersected by the closed form the
//     general formula reduces to, bit for bit (see intersect_axis);
//   * a FRONT_ONLY quad facing away from the ray is skipped before its intersection,
//     which is exact because such a quad can never become the hit (trace.rs:92);
//   * texel gamma decoding is a 256-entry table of the reference's own decoded values,
//     held in LDS.
#pragma once

#include "lw/abi.h"
#include "lw_math.hpp"

namespace lw {

constexpr uint32_t EMPTY = LW_EMPTY;
constexpr uint32_t UNIFORM = LW_UNIFORM;
constexpr uint32_t MAX_STEPS = 65536u;   // trace.rs:31
constexpr uint32_t MAX_EVENTS = 48u;     // integrator.rs:24
constexpr float COPLANAR = 1e-5f;        // trace.rs:73

// QuadInfo bits: an axis quad has eu = su * e_i, ev = sv * e_j with su, sv = +-1 exactly
// and i != j; its normal eu x ev = sn * e_k.
constexpr uint32_t QI_AXIS = 1u << 31;
constexpr uint32_t QI_NEG_U = 1u << 6;
constexpr uint32_t QI_NEG_V = 1u << 7;
constexpr uint32_t QI_NEG_N = 1u << 8;

// Model face words (DevScene::mfaces). A model whose quads are all FRONT_ONLY axis quads, no
// two with the same normal, has MF_FACES set and, for each normal direction f = 2 k + [the
// normal eu x ev is -e_k], the quad's index within the model in bits 3f .. 3f + 2 (MF_NONE when
// the model has no such quad). Such a quad passes the front-only test of a ray (d) only when
// its normal points along d_k, so a ray can meet at most one quad per axis of the model.
constexpr uint32_t MF_FACES = 1u << 31;
constexpr uint32_t MF_NONE = 7u;

// Device copies of the scene buffers (validated on the host before any launch, so every
// index below stays in bounds).
struct DevScene {
  const uint32_t *__restrict__ grid;
  const lw_model *__restrict__ models;
  const float4 *__restrict__ quads;  // 5 float4 per quad, the GpuQuad layout
  const uint32_t *__restrict__ qinfo;
  const uint32_t *__restrict__ atlas;
  const lw_emitter *__restrict__ emitters;
  const uint32_t *__restrict__ egrid;
  const float *__restrict__ sky;
  const float *__restrict__ gamma_lut;  // 256 decoded byte values
  // Per stored brick, which of its 64 sub-blocks of 4 x 4 x 4 voxels hold a non-zero model
  // (bit ((y/4)*4 + z/4)*4 + x/4). A voxel of an empty sub-block is model 0 without a
  // load of its brick word.
  const uint2 *__restrict__ occupancy;
  // Per brick, bit k (k = [+x] | [+y] << 1 | [+z] << 2) is set when the box of bricks from
  // this one to the grid's end toward octant k (this brick included) holds a brick or an
  // overlay that is not EMPTY.
  const uint8_t *__restrict__ octant_boxes;
  // Per model, its face word (MF_* above), 0 for any other model.
  const uint32_t *__restrict__ mfaces;
  // Per stored brick (by its top-level entry) and voxel (local index), for each ray octant k
  // (bits 4k .. 4k + 3, k as DevScene::octant_boxes numbers octants) the side n (at most 8) of
  // the largest cube of voxels from this one toward octant k, clipped to the brick, that holds
  // no model; 0 for a voxel with a model. Null when the scene has no table.
  const uint32_t *__restrict__ cell_ext;
};

struct Hit {
  float t;
  uint32_t quad, texel;
  float s, tt;
};

// buf[i] through a 32-bit byte offset from the buffer's base, so that the load takes the
// base from scalar registers and the offset from one vector register (every scene buffer is
// smaller than 4 GiB, checked on the host).
template <class T>
__device__ __forceinline__ T ld(const T *__restrict__ buf, uint32_t i) {
  return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(buf) + i * (uint32_t)sizeof(T));
}

__device__ __forceinline__ float get4(float4 a, int i) {
  return i == 0 ? a.x : (i == 1 ? a.y : a.z);
}
__device__ __forceinline__ V3 xyz(float4 a) { return V3{a.x, a.y, a.z}; }

// scene.rs:288-294
__device__ __forceinline__ uint32_t texel(const DevScene &s, const lw_params &p, float u, float v) {
  int32_t w = (int32_t)p.off2[1];
  int32_t h = (int32_t)p.off2[2];
  int32_t x = clamp_i(as_i32(f_floor(u)), 0, w - 1);
  int32_t y = clamp_i(as_i32(f_floor(v)), 0, h - 1);
  return ld(s.atlas, (uint32_t)(y * w + x));
}

// integrator.rs:27-34 with scene.rs:299-306: (byte / 255.0) decoded by gamma_to_linear,
// read from the table of exactly those values.
__device__ __forceinline__ V3 linear_rgb(const float *lut, uint32_t c) {
  return V3{lut[c & 0xFFu], lut[(c >> 8) & 0xFFu], lut[(c >> 16) & 0xFFu]};
}

// trace.rs:37-58, with the FRONT_ONLY cull of trace.rs:92 taken first. The cull uses
// q.normal() . d = (ev x eu) . d, which is exactly -(eu x ev) . d = -denom, so "culled"
// is denom <= 0 (a NaN denom culls nothing and intersects nothing, as in the reference).
__device__ __forceinline__ bool intersect_general(const float4 *q, float4 p0, bool front_only,
                                                  V3 o, V3 d, float tmin, float tmax, float &t_out,
                                                  float &s_out, float &tt_out) {
  V3 eu = xyz(q[1]), ev = xyz(q[2]);
  V3 n = cross(eu, ev);
  float denom = dot(n, d);
  if (front_only && denom <= 0.0f) return false;
  if (f_abs(denom) < 1e-12f) return false;
  float t = dot(n, xyz(p0) - o) / denom;
  if (!(t > tmin && t < tmax)) return false;
  V3 p = o + d * t - xyz(p0);
  float nn = dot(n, n);
  float s = dot(cross(p, ev), n) / nn;
  float tt = dot(cross(eu, p), n) / nn;
  if (s < 0.0f || s > 1.0f || tt < 0.0f || tt > 1.0f) return false;
  t_out = t;
  s_out = s;
  tt_out = tt;
  return true;
}

// The general formula for an axis quad (eu = su e_i, ev = sv e_j, n = sn e_k), for
// finite rays: every product with a zero component of n, eu or ev is a zero and every
// product with +-1 is exact, so n.d = sn d_k, n.(p0 - o) = sn (p0_k - o_k), n.n = 1,
// s = su p_i and tt = sv p_j, and t = (p0_k - o_k) / d_k has the same bits as the
// general quotient (a negated numerator over a negated denominator). The zero terms can
// only change the sign of a zero; a zero t is rejected for tmin >= 0, and a zero s or tt
// passes the range test and the texel clamp alike. A zero numerator with tmin < 0 is
// left to the general formula.
__device__ __forceinline__ bool intersect_quad(const float4 *q, uint32_t info, float4 p0,
                                               bool front_only, V3 o, V3 d, float tmin,
                                               float tmax, float &t_out, float &s_out,
                                               float &tt_out) {
  if (!(info & QI_AXIS))
    return intersect_general(q, p0, front_only, o, d, tmin, tmax, t_out, s_out, tt_out);
  const int k = info & 3, i = (info >> 2) & 3, j = (info >> 4) & 3;
  float dk = get(d, k);
  float denom = (info & QI_NEG_N) ? -dk : dk;
  if (front_only && denom <= 0.0f) return false;
  if (f_abs(dk) < 1e-12f) return false;
  float c = get4(p0, k) - get(o, k);
  if (c == 0.0f && !(tmin >= 0.0f))
    return intersect_general(q, p0, front_only, o, d, tmin, tmax, t_out, s_out, tt_out);
  float t = c / dk;
  if (!(t > tmin && t < tmax)) return false;
  float pi = (get(o, i) + get(d, i) * t) - get4(p0, i);
  float pj = (get(o, j) + get(d, j) * t) - get4(p0, j);
  float s = (info & QI_NEG_U) ? -pi : pi;
  float tt = (info & QI_NEG_V) ? -pj : pj;
  if (s < 0.0f || s > 1.0f || tt < 0.0f || tt > 1.0f) return false;
  t_out = t;
  s_out = s;
  tt_out = tt;
  return true;
}

// trace.rs:81-109 (with quad_texel, trace.rs:62-69)
__device__ __forceinline__ void test_quad(const DevScene &s, const lw_params &p, uint32_t qi,
                                          V3 o, V3 d, float tmin, Hit &best) {
  const float4 *q = s.quads + (size_t)qi * 5;
  float4 p0 = ld(s.quads, qi * 5u);
  uint32_t flags = f_bits(p0.w);
  uint32_t info = ld(s.qinfo, qi);
  float reach = (best.quad != UINT32_MAX && qi > best.quad) ? best.t + COPLANAR : best.t;
  float t, ss, tt;
  if (!intersect_quad(q, info, p0, (flags & LW_FLAG_FRONT_ONLY) != 0, o, d, tmin, reach, t, ss,
                      tt))
    return;
  float4 uv = ld(s.quads, qi * 5u + 3u);
  float4 uvt = ld(s.quads, qi * 5u + 4u);
  float cs = f_min(f_max(ss, 0.0005f), 0.9995f);
  float ct = f_min(f_max(tt, 0.0005f), 0.9995f);
  float u = uv.x + cs * uv.z + ct * uvt.x;
  float v = uv.y + cs * uv.w + ct * uvt.y;
  uint32_t tx = texel(s, p, u, v);
  uint32_t alpha = tx >> 24;
  uint32_t mode = flags & LW_FLAG_ALPHA_MASK;
  if ((mode == LW_ALPHA_CUTOUT && alpha < 128u) || (mode == LW_ALPHA_TRANSLUCENT && alpha == 0u))
    return;
  best = Hit{t, qi, tx, ss, tt};
}

// trace.rs:113-116
__device__ __forceinline__ bool is_see_through(const lw_params &p, uint32_t qi) {
  return qi >= p.see_through[0] && qi - p.see_through[0] < p.see_through[1];
}

// trace.rs:120-131
__device__ __forceinline__ bool see_through_hit(const DevScene &s, const lw_params &p, V3 o, V3 d,
                                             float tmin, Hit &out) {
  Hit best{INFINITY, UINT32_MAX, 0u, 0.0f, 0.0f};
  for (uint32_t i = 0; i < p.see_through[1]; i++) test_quad(s, p, p.see_through[0] + i, o, d, tmin, best);
  if (best.quad != UINT32_MAX) {
    out = best;
    return true;
  }
  return false;
}

// trace.rs:293-319
__device__ __forceinline__ float safe_inv(float x) {
  if (f_abs(x) < 1e-12f) return x < 0.0f ? -1e12f : 1e12f;
  return 1.0f / x;
}
__device__ __forceinline__ int32_t sign_i(float x) { return x > 0.0f ? 1 : (x < 0.0f ? -1 : 0); }
__device__ __forceinline__ float next_boundary(float o, float inv, int32_t cell, int32_t step) {
  if (step == 0) return INFINITY;
  float bound = (float)(step > 0 ? cell + 1 : cell);
  return (bound - o) * inv;
}

// The first two arguments of every kernel that traces, (const lw_params p, const DevScene s),
// as its kernel argument segment holds them: each argument at the next offset aligned for its
// type, so s right after p. Every kernel that reaches trace, walk_begin, walk_step or
// sample_event takes these two first (tests/test_library.py checks the sources for it).
struct KernelArgs {
  lw_params p;
  DevScene s;
};
static_assert(offsetof(KernelArgs, s) == sizeof(lw_params), "DevScene follows lw_params unpadded");

// The kernel's (p, s) read again from its argument segment through a pointer the compiler
// cannot trace back to the kernel's entry, so that each field is loaded (a scalar load, from
// the scalar cache) where it is used. Without it, every field the kernel uses anywhere is
// loaded at its entry and held in scalar registers for the whole kernel; at this kernel's size
// most of them are spilled to lanes of a vector register, and each use inside the walk loop
// then costs a vector instruction (v_readlane) to read it back.
__device__ __forceinline__ const KernelArgs &kernel_args() {
  typedef const KernelArgs __attribute__((address_space(4))) *Seg;
  Seg q = (Seg)__builtin_amdgcn_kernarg_segment_ptr();
  asm volatile("" : "+s"(q));
  return *(const KernelArgs *)q;
}

// trace.rs:137-290 as a resumable walk: walk_begin sets up the ray (false: it misses the grid
// or starts beyond tmax) and walk_step runs one iteration of the traversal loop, ending it with
// a hit (WALK_HIT, the hit in `out`) or a miss (WALK_MISS), or leaving it to go on (WALK_ON).
// trace() runs a walk to its end; a caller can also run two walks one after the other in one
// loop. Both read the scene and its layout from kernel_args(). cell_out (when non-null)
// receives the cell the traversal stopped in.
constexpr int WALK_ON = 0, WALK_HIT = 1, WALK_MISS = 2;

struct Walk {
  V3 o, d, inv;
  float tmin, t_end, best_t, t;
  float next[3];
  int32_t cell[3], step[3];
  uint32_t it, cached_b, entry, ov, octant;
  uint2 occ;
};

__device__ __forceinline__ bool walk_begin(V3 o, V3 d, float tmin, float tmax, Walk &w,
                                           int32_t *cell_out) {
  const lw_params &p = kernel_args().p;
  const uint32_t dx = p.grid_dims[0], dy = p.grid_dims[1], dz = p.grid_dims[2];
  const V3 size = V3{(float)(dx * 16u), (float)(dy * 16u), (float)(dz * 16u)};

  const V3 inv = v3(safe_inv(d.x), safe_inv(d.y), safe_inv(d.z));
  const V3 ta = mul(V3{0.0f, 0.0f, 0.0f} - o, inv);
  const V3 tb = mul(size - o, inv);
  const float t_enter = f_max(f_max(f_min(ta.x, tb.x), f_min(ta.y, tb.y)), f_min(ta.z, tb.z));
  const float t_exit = f_min(f_min(f_max(ta.x, tb.x), f_max(ta.y, tb.y)), f_max(ta.z, tb.z));
  const float t = f_max(t_enter, 0.0f);
  const float t_end = f_min(t_exit, tmax);
  if (cell_out) cell_out[0] = cell_out[1] = cell_out[2] = -1;
  if (t > t_end) return false;

  const V3 start = o + d * (t + 1e-4f);
  const int32_t gs[3] = {(int32_t)dx * 16, (int32_t)dy * 16, (int32_t)dz * 16};
  w.cell[0] = clamp_i(as_i32(f_floor(start.x)), 0, gs[0] - 1);
  w.cell[1] = clamp_i(as_i32(f_floor(start.y)), 0, gs[1] - 1);
  w.cell[2] = clamp_i(as_i32(f_floor(start.z)), 0, gs[2] - 1);
  w.step[0] = sign_i(d.x);
  w.step[1] = sign_i(d.y);
  w.step[2] = sign_i(d.z);
#pragma unroll
  for (int a = 0; a < 3; a++) w.next[a] = next_boundary(get(o, a), get(inv, a), w.cell[a], w.step[a]);
  w.o = o;
  w.d = d;
  w.inv = inv;
  w.tmin = tmin;
  w.t_end = t_end;
  w.best_t = t_end + 1e-3f;
  w.t = t;
  w.it = 0;
  w.cached_b = UINT32_MAX;
  w.entry = EMPTY;
  w.ov = EMPTY;
  w.octant = (w.step[0] >= 0 ? 1u : 0u) | (w.step[1] >= 0 ? 2u : 0u) | (w.step[2] >= 0 ? 4u : 0u);
  w.occ = make_uint2(0u, 0u);
  return true;
}

// The loop's increment and bound: MAX_STEPS iterations end the walk without a hit.
__device__ __forceinline__ int walk_next(Walk &w) {
  w.it += 1u;
  return w.it < MAX_STEPS ? WALK_ON : WALK_MISS;
}

__device__ __forceinline__ int walk_step(Walk &w, Hit &out, int32_t *cell_out) {
  const KernelArgs &ka = kernel_args();
  const DevScene &s = ka.s;
  const lw_params &p = ka.p;
  const uint32_t dx = p.grid_dims[0], dz = p.grid_dims[2];
  const int32_t gs[3] = {(int32_t)dx * 16, (int32_t)p.grid_dims[1] * 16, (int32_t)dz * 16};
  const uint32_t *top = s.grid + p.off[0];
  const uint32_t *ov_top = s.grid + p.off[2];
  const V3 o = w.o, d = w.d, inv = w.inv;
  const float tmin = w.tmin, t_end = w.t_end, best_t = w.best_t;
  const int32_t *step = w.step;
  const uint32_t octant = w.octant;
  int32_t *cell = w.cell;
  float *next = w.next;
  float &t = w.t;
  uint32_t &it = w.it, &cached_b = w.cached_b, &entry = w.entry, &ov = w.ov;
  uint2 &occ = w.occ;
  const uint32_t bx = (uint32_t)(cell[0] >> 4);
  const uint32_t by = (uint32_t)(cell[1] >> 4);
  const uint32_t bz = (uint32_t)(cell[2] >> 4);
  const uint32_t b = (by * dz + bz) * dx + bx;
  if (b != cached_b) {
    entry = ld(top, b);
    ov = ld(ov_top, b);
    cached_b = b;
    if (entry != EMPTY && (entry & UNIFORM) == 0) occ = ld(s.occupancy, entry);
  }
  if (cell_out) {
    cell_out[0] = cell[0];
    cell_out[1] = cell[1];
    cell_out[2] = cell[2];
  }

  if (entry == EMPTY && ov == EMPTY) {
    // Each brick coordinate of the walk only ever moves in its axis's step direction (a
    // cell step moves by the step, a brick skip crosses one face in the step's direction
    // and clamps the other coordinates into the brick, a sub-block jump moves cells by the
    // step), and not at all on an axis with step 0. So every later brick lies in the box
    // toward the ray's octant (step 0 counted as +). When that box holds only EMPTY bricks
    // without overlays, no quad will be tested and the result is "no hit" wherever the
    // walk would end (t_end, the grid's faces or MAX_STEPS). Only the final cell differs,
    // so a caller asking for it (cell_out) gets the whole walk.
    if (cell_out == nullptr && ((ld(s.octant_boxes, b) >> octant) & 1u) == 0u) return WALK_MISS;
    // An empty brick is crossed in one move, to the face where the ray leaves it.
    const int32_t bc[3] = {(int32_t)bx, (int32_t)by, (int32_t)bz};
    float t_out = INFINITY;
    int axis = 0;
#pragma unroll
    for (int a = 0; a < 3; a++) {
      if (step[a] != 0) {
        int32_t bound = step[a] > 0 ? (bc[a] + 1) * 16 : bc[a] * 16;
        float tt = ((float)bound - get(o, a)) * get(inv, a);
        if (tt < t_out) {
          t_out = tt;
          axis = a;
        }
      }
    }
    if (t_out > t_end) return WALK_MISS;
    const V3 pos = o + d * t_out;
#pragma unroll
    for (int a = 0; a < 3; a++) {
      if (a == axis) {
        cell[a] = step[a] > 0 ? (bc[a] + 1) * 16 : bc[a] * 16 - 1;
      } else {
        cell[a] = clamp_i(as_i32(f_floor(get(pos, a))), bc[a] * 16, bc[a] * 16 + 15);
      }
    }
    const int32_t ca = axis == 0 ? cell[0] : (axis == 1 ? cell[1] : cell[2]);
    const int32_t ga = axis == 0 ? gs[0] : (axis == 1 ? gs[1] : gs[2]);
    if (ca < 0 || ca >= ga) return WALK_MISS;
#pragma unroll
    for (int a = 0; a < 3; a++) next[a] = next_boundary(get(o, a), get(inv, a), cell[a], step[a]);
    t = t_out;
    return walk_next(w);
  }

  const bool stored = entry != EMPTY && (entry & UNIFORM) == 0;
  // The occupancy bit of the cell's 4^3 sub-block (1 outside stored bricks).
  uint32_t sb_full = 1u;
  if (stored) {
    const uint32_t sb = ((((uint32_t)cell[1] & 15u) >> 2) * 4u + (((uint32_t)cell[2] & 15u) >> 2)) * 4u +
                        (((uint32_t)cell[0] & 15u) >> 2);
    sb_full = ((sb < 32u ? occ.x : occ.y) >> (sb & 31u)) & 1u;
  }
  uint32_t model;
  const uint32_t local =
      (uint32_t)((((cell[1] & 15) * 16 + (cell[2] & 15)) * 16) + (cell[0] & 15));
  if (entry == EMPTY) {
    model = 0;
  } else if ((entry & UNIFORM) != 0) {
    model = entry & ~UNIFORM;
  } else if (sb_full) {
    const uint32_t w = ld(s.grid, p.off[1] + entry * LW_BRICK_WORDS + (local >> 1));
    model = (w >> ((local & 1u) * 16u)) & 0xFFFFu;
  } else {
    model = 0;
  }
  if (stored && ov == EMPTY) {
    // A box of empty cells around the current one, each with no model (and the brick has no
    // overlay): cross it in one move, landing where the cell-by-cell walk below would, with
    // the same t, cells and step count. The walk is a merge of the three axes' crossing times
    // (each a pure function of the cell, next_boundary), taking the smallest and, on a tie,
    // the higher axis. It leaves the box at the first crossing of a far face, ts on axis x;
    // by then each other axis has made the crossings that come before (ts, x) in that order.
    // Taken only when that exit stays in the grid, within t_end and within MAX_STEPS;
    // otherwise the cells are walked one by one. The box is the cell's empty 4^3 sub-block,
    // or the empty 8^3 block holding it (aligned boxes inside the brick), or, for an empty
    // cell of an occupied sub-block, the empty cube from the cell toward the ray's octant
    // (cell_ext), clipped to the brick.
    int32_t far[3];
    bool boxed = false;
    if (sb_full == 0u) {
      // The 8^3 block's eight sub-blocks are the occupancy bits 0x330033 shifted by
      // 2 x8 + 8 z8 + 32 y8.
      const uint32_t x8 = ((uint32_t)cell[0] & 15u) >> 3, y8 = ((uint32_t)cell[1] & 15u) >> 3,
                     z8 = ((uint32_t)cell[2] & 15u) >> 3;
      const int32_t span = ((y8 ? occ.y : occ.x) & (0x330033u << (2u * x8 + 8u * z8))) == 0u ? 8 : 4;
#pragma unroll
      for (int a = 0; a < 3; a++) far[a] = (cell[a] & ~(span - 1)) + (step[a] > 0 ? span - 1 : 0);
      boxed = true;
    } else if (model == 0u && s.cell_ext != nullptr) {
      const int32_t n = (int32_t)((ld(s.cell_ext, entry * 4096u + local) >> (4u * octant)) & 15u);
      if (n >= 2) {
#pragma unroll
        for (int a = 0; a < 3; a++) {
          const int32_t lo = cell[a] & ~15;
          far[a] = step[a] > 0 ? min(cell[a] + n - 1, lo + 15) : (step[a] < 0 ? max(cell[a] - n + 1, lo) : cell[a]);
        }
        boxed = true;
      }
    }
    if (boxed) {
      float tf[3];
#pragma unroll
      for (int a = 0; a < 3; a++) tf[a] = next_boundary(get(o, a), get(inv, a), far[a], step[a]);
      int x;
      if (tf[0] < tf[1]) {
        x = tf[0] < tf[2] ? 0 : 2;
      } else if (tf[1] < tf[2]) {
        x = 1;
      } else {
        x = 2;
      }
      const float ts = x == 0 ? tf[0] : (x == 1 ? tf[1] : tf[2]);
      int32_t nc[3];
      uint32_t moves = 1;
#pragma unroll
      for (int a = 0; a < 3; a++) {
        int32_t c = cell[a];
        if (a == x) {
          moves += (uint32_t)((far[a] - c) * step[a]);
          c = far[a] + step[a];
        } else if (step[a] != 0) {
          // The walk crosses axis a's boundaries from c toward far[a] while each comes
          // before (ts, x). Their times never fall from one crossing to the next (f32
          // rounding is monotone), so the crossings made are a prefix, and its length is
          // found by binary lifting over the at most 7 of them, branch-free.
          const int32_t avail = (far[a] - c) * step[a];
          int32_t k = 0;
#pragma unroll
          for (int32_t st = 4; st >= 1; st >>= 1) {
            const int32_t j = k + st;
            const float tb = next_boundary(get(o, a), get(inv, a), c + (j - 1) * step[a], step[a]);
            if (j <= avail && (tb < ts || (tb == ts && a > x))) k = j;
          }
          c += k * step[a];
          moves += (uint32_t)k;
        }
        nc[a] = c;
      }
      const int32_t cx = x == 0 ? nc[0] : (x == 1 ? nc[1] : nc[2]);
      const int32_t gx = x == 0 ? gs[0] : (x == 1 ? gs[1] : gs[2]);
      if (ts <= t_end && cx >= 0 && cx < gx && it + moves - 1u < MAX_STEPS) {
#pragma unroll
        for (int a = 0; a < 3; a++) {
          cell[a] = nc[a];
          next[a] = next_boundary(get(o, a), get(inv, a), cell[a], step[a]);
        }
        t = ts;
        it += moves - 1u;
        return walk_next(w);
      }
    }
  }

  const float t_cell_exit = f_min(next[0], f_min(next[1], next[2]));
  const float t_lo = f_max(t - 1e-4f, tmin);
  const float t_hi = t_cell_exit + 1e-4f;
  Hit cand{f_min(t_hi, best_t), UINT32_MAX, 0u, 0.0f, 0.0f};
  if (model != 0) {
    const lw_model m = ld(s.models, model);
    const uint32_t fw = ld(s.mfaces, model);
    const V3 lo = o - v3((float)cell[0], (float)cell[1], (float)cell[2]);
    // A face model's quads that can pass this ray's front-only test, one per axis with a
    // non-zero step (the quad whose normal has the step's sign), tested in quad order. The
    // others would be culled by test_quad without a candidate, so skipping them leaves the
    // same best hit: a wave of cube visits runs 3 quad tests instead of 6.
    uint32_t n = m.quad_count, order = 0u;
    if (fw & MF_FACES) {
      uint32_t c[3];
#pragma unroll
      for (int a = 0; a < 3; a++)
        c[a] = step[a] == 0 ? MF_NONE : (fw >> (3u * (2u * (uint32_t)a + (step[a] < 0 ? 1u : 0u)))) & 7u;
      const uint32_t s0 = min(min(c[0], c[1]), c[2]), s2 = max(max(c[0], c[1]), c[2]);
      const uint32_t s1 = c[0] ^ c[1] ^ c[2] ^ s0 ^ s2;
      n = (c[0] != MF_NONE ? 1u : 0u) + (c[1] != MF_NONE ? 1u : 0u) + (c[2] != MF_NONE ? 1u : 0u);
      order = s0 | (s1 << 3) | (s2 << 6);
    }
    for (uint32_t i = 0; i < n; i++) {
      const uint32_t k = (fw & MF_FACES) ? (order >> (3u * i)) & 7u : i;
      test_quad(s, p, m.quad_offset + k, lo, d, t_lo, cand);
    }
  }
  if (ov != EMPTY) {
    const uint32_t base = p.off[3] + ov * LW_OV_CELL_WORDS + local * 2u;
    const uint32_t off = ld(s.grid, base);
    const uint32_t count = ld(s.grid, base + 1);
    for (uint32_t i = 0; i < count; i++) {
      const uint32_t qi = ld(s.grid, p.off2[0] + off + i);
      test_quad(s, p, qi, o, d, t_lo, cand);
    }
  }
  if (cand.quad != UINT32_MAX) {
    out = cand;
    return WALK_HIT;
  }

  int axis;
  if (next[0] < next[1]) {
    axis = next[0] < next[2] ? 0 : 2;
  } else if (next[1] < next[2]) {
    axis = 1;
  } else {
    axis = 2;
  }
  t = axis == 0 ? next[0] : (axis == 1 ? next[1] : next[2]);
  if (t > t_end) return WALK_MISS;
  bool out_of_grid = false;
#pragma unroll
  for (int a = 0; a < 3; a++) {
    if (a == axis) {
      cell[a] += step[a];
      out_of_grid = cell[a] < 0 || cell[a] >= gs[a];
      next[a] = next_boundary(get(o, a), get(inv, a), cell[a], step[a]);
    }
  }
  if (out_of_grid) return WALK_MISS;
  return walk_next(w);
}

// s and p are the kernel's own first two arguments (the walk reads them from kernel_args()).
__device__ __forceinline__ bool trace(const DevScene &s, const lw_params &p, V3 o, V3 d, float tmin,
                                   float tmax, Hit &out, int32_t *cell_out) {
  Walk w;
  if (!walk_begin(o, d, tmin, tmax, w, cell_out)) return false;
  for (;;) {
    const int r = walk_step(w, out, cell_out);
    if (r != WALK_ON) return r == WALK_HIT;
  }
}


// ---- integrator.rs ----

// integrator.rs:107-127
__device__ __forceinline__ V3 rgb_from_hsl(float hue, float saturation, float lightness) {
  float c = f_min(1.0f, (1.0f - f_abs(2.0f * lightness - 1.0f)) * saturation);
  float h = hue * 6.0f;
  float x = c * (1.0f - f_abs(rem2(h) - 1.0f));
  V3 rgb;
  if (h < 1.0f) {
    rgb = v3(c, x, 0.0f);
  } else if (h < 2.0f) {
    rgb = v3(x, c, 0.0f);
  } else if (h < 3.0f) {
    rgb = v3(0.0f, c, x);
  } else if (h < 4.0f) {
    rgb = v3(0.0f, x, c);
  } else if (h < 5.0f) {
    rgb = v3(x, 0.0f, c);
  } else {
    rgb = v3(c, 0.0f, x);
  }
  float m = f_max(0.0f, lightness - 0.5f * c);
  return rgb + v3(m, m, m);
}

// integrator.rs:158-188
__device__ __forceinline__ V3 sky_radiance(const DevScene &s, const lw_params &p, V3 d) {
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
  const float *a00 = s.sky + (i0 * m + j0) * 3u;
  const float *a10 = s.sky + (i1 * m + j0) * 3u;
  const float *a01 = s.sky + (i0 * m + j1) * 3u;
  const float *a11 = s.sky + (i1 * m + j1) * 3u;
  float hsl[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    float y0 = a00[k] + (x - fx) * (a10[k] - a00[k]);
    float y1 = a01[k] + (x - fx) * (a11[k] - a01[k]);
    hsl[k] = y0 + (y - fy) * (y1 - y0);
  }
  return rgb_from_hsl(hsl[0], hsl[1], hsl[2]) * p.sky_sun[3];
}

// integrator.rs:196-217
__device__ __forceinline__ V3 map_lookup(const DevScene &s, uint32_t base, float w, float h, V3 d) {
  float u = f_atan2(d.z, d.x) / (2.0f * PI_F);
  u = rem1(rem1(u) + 1.0f);
  const float v = 1.0f - f_acos(f_clamp(d.y, -1.0f, 1.0f)) / PI_F;
  const float x = u * w - 0.5f;
  const float y = f_clamp(v * h - 0.5f, 0.0f, h - 1.0f);
  const float fx = f_floor(x);
  const float fy = f_floor(y);
  const float tx = x - fx, ty = y - fy;
  const float i0 = fx < 0.0f ? w - 1.0f : fx;
  const float i1 = fx + 1.0f >= w ? 0.0f : fx + 1.0f;
  const float j1 = f_min(fy + 1.0f, h - 1.0f);
  const uint32_t wu = as_u32(w);
  auto at = [&](float i, float j) -> V3 {
    const float *k = s.sky + (base + (as_u32(j) * wu + as_u32(i)) * 3u);
    return V3{k[0], k[1], k[2]};
  };
  const V3 top = at(i0, fy) * (1.0f - tx) + at(i1, fy) * tx;
  const V3 bottom = at(i0, j1) * (1.0f - tx) + at(i1, j1) * tx;
  return top * (1.0f - ty) + bottom * ty;
}

// integrator.rs:225-231
__device__ __forceinline__ V3 sky_lens(const DevScene &s, const lw_params &p, V3 d) {
  if (p.sky_map[0] < 1.0f) return sky_radiance(s, p, d);
  return map_lookup(s, 0u, p.sky_map[0], p.sky_map[1], d);
}

// integrator.rs:237-243
__device__ __forceinline__ V3 sky_light(const DevScene &s, const lw_params &p, V3 d) {
  if (p.sky_map[0] < 1.0f) return sky_radiance(s, p, d);
  return map_lookup(s, as_u32(p.sky_map[0] * p.sky_map[1]) * 3u, p.sky_map[2], p.sky_map[3], d);
}

// integrator.rs:252-277
__device__ __forceinline__ V3 sun_texture(const DevScene &s, const lw_params &p, const float *lut,
                                       V3 d) {
  const float *sprite = p.sun_sprite;
  if (sprite[2] <= 0.0f || dot(d, V3{p.sun_dir[0], p.sun_dir[1], p.sun_dir[2]}) < 0.5f)
    return V3{0.0f, 0.0f, 0.0f};
  const float width = p.sun_sv[3];
  const float width2 = width * 2.0f;
  const float a = 0.5f * PI_F - f_acos(dot(d, V3{p.sun_su[0], p.sun_su[1], p.sun_su[2]})) + width;
  const float b = 0.5f * PI_F - f_acos(dot(d, V3{p.sun_sv[0], p.sun_sv[1], p.sun_sv[2]})) + width;
  if (a < 0.0f || a >= width2 || b < 0.0f || b >= width2) return V3{0.0f, 0.0f, 0.0f};
  const float x = f_clamp(f_floor(a / width2 * sprite[2] - 5e-8f), 0.0f, sprite[2] - 1.0f);
  const float y = f_clamp(f_floor((1.0f - b / width2) * sprite[3] - 5e-8f), 0.0f, sprite[3] - 1.0f);
  return linear_rgb(lut, texel(s, p, sprite[0] + x + 0.5f, sprite[1] + y + 0.5f)) * p.sun_su[3];
}

// integrator.rs:279-287
__device__ __forceinline__ V3 cosine_hemisphere(V3 n, Rng &rng) {
  const float r1 = rng.next();
  const float r2 = rng.next();
  const float r = f_sqrt(r1);
  float sin_phi, cos_phi;
  sin_cos_turns(r2, sin_phi, cos_phi);
  V3 t, b;
  onb(n, t, b);
  return normalize(t * (r * cos_phi) + b * (r * sin_phi) + n * f_sqrt(f_max(1.0f - r1, 0.0f)));
}

// integrator.rs:289-302
__device__ __forceinline__ V3 sample_sun(const lw_params &p, Rng &rng) {
  const V3 sun = V3{p.sun_dir[0], p.sun_dir[1], p.sun_dir[2]};
  const float cos_max = p.sun_dir[3];
  const float r1 = rng.next();
  const float r2 = rng.next();
  const float cos_t = 1.0f - r1 * (1.0f - cos_max);
  const float sin_t = f_sqrt(f_max(1.0f - cos_t * cos_t, 0.0f));
  float sin_phi, cos_phi;
  sin_cos_turns(r2, sin_phi, cos_phi);
  V3 t, b;
  onb(sun, t, b);
  return normalize(t * (sin_t * cos_phi) + b * (sin_t * sin_phi) + sun * cos_t);
}

__device__ __forceinline__ V3 quad_normal(const float4 *q) { return cross(xyz(q[2]), xyz(q[1])); }

// integrator.rs:318-346
// n_traces, when given, counts the trace calls (the device's ray count).
__device__ __forceinline__ V3 shadow(const DevScene &s, const lw_params &p, const float *lut, V3 o,
                                  V3 d, float tmax, uint32_t *n_traces = nullptr) {
  const float eps = p.misc_f[1];
  V3 tr{1.0f, 1.0f, 1.0f};
  for (int i = 0; i < 8; i++) {
    Hit h;
    if (n_traces) ++*n_traces;
    if (!trace(s, p, o, d, eps, tmax, h, nullptr)) return tr;
    const float4 *q = s.quads + (size_t)h.quad * 5;
    const uint32_t flags = f_bits(q[0].w);
    if ((flags & LW_FLAG_WATER) != 0) {
      if (dot(quad_normal(q), d) > 0.0f)
        tr = mul(tr, transmittance(V3{p.water_color[0], p.water_color[1], p.water_color[2]}, h.t));
    } else if ((flags & LW_FLAG_ALPHA_MASK) == LW_ALPHA_TRANSLUCENT) {
      const float a = (float)(h.texel >> 24) / 255.0f;
      const V3 col = mul(linear_rgb(lut, h.texel), linear_rgb(lut, f_bits(q[1].w)));
      tr = mul(tr, lerp(V3{1.0f, 1.0f, 1.0f}, col, a));
    } else {
      return V3{0.0f, 0.0f, 0.0f};
    }
    o = o + d * (h.t + eps);
    tmax -= h.t + eps;
  }
  return tr;
}

// integrator.rs:349-353
__device__ __forceinline__ uint32_t pick(uint32_t n, float u) {
  const uint32_t i = as_u32(u * (float)n);
  return i < n ? i : n - 1;
}

// integrator.rs:357-373
__device__ __forceinline__ void emitter_cell(const DevScene &s, V3 p, uint32_t &start,
                                             uint32_t &count) {
  const uint32_t *g = s.egrid;
  const float cell = (float)g[0];
  const int32_t sx = (int32_t)g[1], sy = (int32_t)g[2], sz = (int32_t)g[3];
  const int32_t cx = (int32_t)((uint32_t)as_i32(f_floor(p.x / cell)) - g[4]);
  const int32_t cy = (int32_t)((uint32_t)as_i32(f_floor(p.y / cell)) - g[5]);
  const int32_t cz = (int32_t)((uint32_t)as_i32(f_floor(p.z / cell)) - g[6]);
  if (cx < 0 || cy < 0 || cz < 0 || cx >= sx || cy >= sy || cz >= sz) {
    start = 0;
    count = 0;
    return;
  }
  const uint32_t i = ((uint32_t)cy * (uint32_t)sz + (uint32_t)cz) * (uint32_t)sx + (uint32_t)cx;
  start = g[LW_EMITTER_GRID_HEADER + i * 2u];
  count = g[LW_EMITTER_GRID_HEADER + i * 2u + 1u];
}

// integrator.rs:380-414
__device__ __forceinline__ V3 sample_emitter_face(const DevScene &s, const lw_params &p,
                                                  const float *lut, V3 origin, V3 n,
                                                  const lw_emitter &e, uint32_t face, float scaler,
                                                  Rng &rng, uint32_t *n_traces = nullptr) {
  const float4 *q = s.quads + (size_t)(f_bits(e.pos[3]) + face) * 5;
  const float ss = rng.next();
  const float t = rng.next();
  const float4 q0 = q[0], q1 = q[1], q2 = q[2];
  const V3 at = V3{e.pos[0], e.pos[1], e.pos[2]} + xyz(q0) + xyz(q1) * ss + xyz(q2) * t;
  const V3 to = at - origin;
  if (dot(to, n) <= 0.0f) return V3{0.0f, 0.0f, 0.0f};
  const float dist = length(to);
  const V3 l = to / dist;
  const float eps = p.misc_f[1];
  Hit h;
  if (n_traces) ++*n_traces;
  if (!trace(s, p, origin, l, eps, dist + 2.0f * eps, h, nullptr)) return V3{0.0f, 0.0f, 0.0f};
  if (f_abs(h.t - dist) >= 2.0f * eps) return V3{0.0f, 0.0f, 0.0f};
  const V3 nq = cross(xyz(q2), xyz(q1));
  const float area = length(nq);
  const float cos_l = f_abs(dot(nq, l)) / area;
  const float4 *hq = s.quads + (size_t)h.quad * 5;
  const V3 c = mul(linear_rgb(lut, h.texel), linear_rgb(lut, f_bits(hq[1].w)));
  return c * (cos_l / f_max(dist * dist, 1.0f) * area * q2.w * p.misc_f[0] * scaler);
}

// integrator.rs:418-433
__device__ __forceinline__ V3 sample_emitter(const DevScene &s, const lw_params &p, const float *lut,
                                          V3 origin, V3 n, const lw_emitter &e, Rng &rng,
                                          uint32_t *n_traces = nullptr) {
  const uint32_t faces = e.faces[0];
  if (faces == 0) return V3{0.0f, 0.0f, 0.0f};
  if (p.emitter[0] == LW_EMITTER_SAMPLING_ONE) {
    const uint32_t f = pick(faces, rng.next());
    return sample_emitter_face(s, p, lut, origin, n, e, f, 1.0f, rng, n_traces);
  }
  V3 sum{0.0f, 0.0f, 0.0f};
  for (uint32_t f = 0; f < faces; f++)
    sum = sum + sample_emitter_face(s, p, lut, origin, n, e, f, 1.0f / (float)faces, rng, n_traces);
  return sum;
}

// integrator.rs:441-469
__device__ __forceinline__ V3 sample_emitters(const DevScene &s, const lw_params &p,
                                              const float *lut, V3 pos, V3 origin, V3 n, Rng &rng,
                                              uint32_t *n_traces = nullptr) {
  uint32_t start, count;
  emitter_cell(s, pos, start, count);
  if (count == 0) return V3{0.0f, 0.0f, 0.0f};
  const uint32_t *g = s.egrid;
  const size_t list = (size_t)LW_EMITTER_GRID_HEADER + (size_t)g[1] * (size_t)g[2] * (size_t)g[3] * 2 +
                      (size_t)start;
  const uint32_t strategy = p.emitter[0];
  if (strategy == LW_EMITTER_SAMPLING_ONE || strategy == LW_EMITTER_SAMPLING_ONE_BLOCK) {
    const uint32_t i = g[list + pick(count, rng.next())];
    const lw_emitter e = s.emitters[i];
    return sample_emitter(s, p, lut, origin, n, e, rng, n_traces) * PI_F;
  }
  if (strategy == LW_EMITTER_SAMPLING_ALL) {
    V3 sum{0.0f, 0.0f, 0.0f};
    for (uint32_t k = 0; k < count; k++) {
      const lw_emitter e = s.emitters[g[list + k]];
      sum = sum + sample_emitter(s, p, lut, origin, n, e, rng, n_traces);
    }
    return sum * (PI_F / (float)count);
  }
  return V3{0.0f, 0.0f, 0.0f};
}

// integrator.rs:477-505
__device__ __forceinline__ void water_interface(const lw_params &p, V3 pos, V3 d, V3 n, bool front,
                                                bool in_water, float eps, Rng &rng, V3 &d_out,
                                                V3 &o_out, bool &w_out) {
  const float ior = p.water_color[3];
  const float eta = front ? 1.0f / ior : ior;
  const float cos_i = -dot(n, d);
  const float sin2_t = eta * eta * f_max(1.0f - cos_i * cos_i, 0.0f);
  float r0 = (1.0f - ior) / (1.0f + ior);
  r0 = r0 * r0;
  const float m = 1.0f - cos_i;
  const float fresnel = r0 + (1.0f - r0) * m * m * m * m * m;
  if (sin2_t >= 1.0f || rng.next() < fresnel) {
    d_out = normalize(d + n * (2.0f * cos_i));
    o_out = pos + n * eps;
    w_out = in_water;
  } else {
    const float cos_t = f_sqrt(1.0f - sin2_t);
    d_out = normalize(d * eta + n * (eta * cos_i - cos_t));
    o_out = pos - n * eps;
    w_out = front;
  }
}

// integrator.rs:510-526
__device__ __forceinline__ bool water_plane(const lw_params &p, V3 o, V3 d, float &t_out) {
  if (p.water_plane[1] == 0.0f || d.y == 0.0f) return false;
  const float t = (p.water_plane[0] - o.y) / d.y;
  if (!(t > 0.0f && t < INFINITY)) return false;
  const V3 a = o + d * t;
  const float sx = (float)(p.grid_dims[0] * 16u), sz = (float)(p.grid_dims[2] * 16u);
  if (a.x >= 0.0f && a.z >= 0.0f && a.x < sx && a.z < sz) return false;
  t_out = t;
  return true;
}

// integrator.rs:530-534
__device__ __forceinline__ bool under_water_plane(const lw_params &p, V3 o, V3 d) {
  return p.water_plane[1] != 0.0f && o.y < p.water_plane[0] && d.y <= 0.0f;
}

// The camera of one view of a batch: what render_sample reads from Params to make its
// camera ray (integrator.rs:541-556), and the view's sample count.
struct ViewCam {
  float cam_pos[4], cam_fwd[4], cam_right[4], cam_up[4];
  uint32_t width, height, seed, spp;
};
__device__ __forceinline__ uint32_t cam_width(const lw_params &c) { return c.frame[0]; }
__device__ __forceinline__ uint32_t cam_height(const lw_params &c) { return c.frame[1]; }
__device__ __forceinline__ uint32_t cam_seed(const lw_params &c) { return c.misc_u[0]; }
__device__ __forceinline__ uint32_t cam_width(const ViewCam &c) { return c.width; }
__device__ __forceinline__ uint32_t cam_height(const ViewCam &c) { return c.height; }
__device__ __forceinline__ uint32_t cam_seed(const ViewCam &c) { return c.seed; }

// render_sample (integrator.rs:539-709) in two parts: sample_begin makes the camera ray
// (:540-569), sample_event runs one iteration of the event loop (:571-707) and says whether
// the path goes on. render_sample runs them to the end; a kernel can also interleave paths
// (start the next sample of a lane as soon as its path ends) with the same arithmetic.
struct SampleState {
  V3 o, d, radiance, beta;
  Rng rng;
  float tmin;
  uint32_t bounces, ev;
  bool specular, in_water;
};

template <class Cam>
__device__ __forceinline__ void sample_begin(SampleState &S, const Cam &c, uint32_t px, uint32_t py,
                                             uint32_t sample) {
  const uint32_t width = cam_width(c);
  const uint32_t height = cam_height(c);
  S.rng = Rng::make(py * width + px, sample, cam_seed(c));
  const float tan_half = c.cam_pos[3];
  const float aspect = c.cam_fwd[3];
  const float jx = S.rng.next();
  const float jy = S.rng.next();
  const float sx = (2.0f * ((float)px + jx) / (float)width - 1.0f) * aspect * tan_half;
  const float sy = (1.0f - 2.0f * ((float)py + jy) / (float)height) * tan_half;
  S.d = normalize(V3{c.cam_fwd[0], c.cam_fwd[1], c.cam_fwd[2]} +
                  V3{c.cam_right[0], c.cam_right[1], c.cam_right[2]} * sx +
                  V3{c.cam_up[0], c.cam_up[1], c.cam_up[2]} * sy);
  S.o = V3{c.cam_pos[0], c.cam_pos[1], c.cam_pos[2]};
  S.radiance = V3{0.0f, 0.0f, 0.0f};
  S.beta = V3{1.0f, 1.0f, 1.0f};
  S.bounces = 0;
  S.ev = 0;
  S.specular = true;
  S.in_water = false;
  S.tmin = 0.0f;
}

// One iteration of the event loop; false once the path has ended (a `break` of the loop or
// MAX_EVENTS iterations), true where the loop would go on (`continue` or the loop's end).
// s_in and p_in are the kernel's own first two arguments.
__device__ __forceinline__ bool sample_event(SampleState &S, const DevScene &s_in,
                                             const lw_params &p_in, const float *lut,
                                             uint32_t *n_traces = nullptr) {
  if (S.ev >= MAX_EVENTS) return false;
  S.ev++;
  Hit hit;
  if (n_traces) ++*n_traces;
  bool has_hit = trace(s_in, p_in, S.o, S.d, S.tmin, INFINITY, hit, nullptr);
  // The rest of the event reads the kernel's (p, s) afresh (kernel_args).
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
  if (S.bounces == 0 && p.see_through[1] > 0 && has_hit && !is_see_through(p, hit.quad)) {
    Hit st;
    if (see_through_hit(s, p, S.o, S.d, hit.t, st)) {
      const uint32_t sflags = f_bits(s.quads[(size_t)st.quad * 5].w);
      if ((sflags & LW_FLAG_ALPHA_MASK) == LW_ALPHA_TRANSLUCENT) {
        if (S.rng.next() < (float)(st.texel >> 24) / 255.0f) {
          hit = st;
          forced = true;
        }
      } else {
        hit = st;
      }
    }
  }
  float tp;
  if (water_plane(p, S.o, S.d, tp) && (!has_hit || tp < hit.t)) {
    if (S.in_water) S.beta = mul(S.beta, transmittance(sigma, tp));
    const bool front = S.d.y < 0.0f;
    const V3 n = v3(0.0f, front ? 1.0f : -1.0f, 0.0f);
    water_interface(p, S.o + S.d * tp, S.d, n, front, S.in_water, eps, S.rng, S.d, S.o, S.in_water);
    S.specular = true;
    S.tmin = 0.0f;
    return true;
  }
  if (!has_hit) {
    if (!S.in_water && !under_water_plane(p, S.o, S.d)) {
      if (S.bounces == 0) {
        S.radiance = S.radiance + mul(S.beta, sky_lens(s, p, S.d));
      } else {
        S.radiance = S.radiance + mul(S.beta, sky_light(s, p, S.d));
      }
      if (S.specular && S.bounces == 0) S.radiance = S.radiance + mul(S.beta, sun_texture(s, p, lut, S.d));
    }
    return false;
  }
  if (S.in_water) S.beta = mul(S.beta, transmittance(sigma, hit.t));
  const float4 *q = s.quads + (size_t)hit.quad * 5;
  const float4 q0 = q[0], q1 = q[1], q2 = q[2];
  const uint32_t flags = f_bits(q0.w);
  const V3 pos = S.o + S.d * hit.t;
  const float alpha = (float)(hit.texel >> 24) / 255.0f;
  const V3 albedo = mul(linear_rgb(lut, hit.texel), linear_rgb(lut, f_bits(q1.w)));
  // integrator.rs:307-315 (shading_normal)
  V3 n = normalize(cross(xyz(q2), xyz(q1)));
  const bool front = dot(n, S.d) < 0.0f;
  if (!front) n = -n;

  if ((flags & LW_FLAG_WATER) != 0) {
    water_interface(p, pos, S.d, n, front, S.in_water, eps, S.rng, S.d, S.o, S.in_water);
    S.specular = true;
    S.tmin = 0.0f;
    return true;
  }

  if (!forced && (flags & LW_FLAG_ALPHA_MASK) == LW_ALPHA_TRANSLUCENT && S.rng.next() >= alpha) {
    S.beta = mul(S.beta, lerp(V3{1.0f, 1.0f, 1.0f}, albedo, alpha));
    S.o = pos + S.d * eps;
    S.tmin = 0.0f;
    return true;
  }

  if (S.bounces >= max_bounces) return false;
  const float emittance = q2.w * p.misc_f[0];
  const bool emits = emittance > 0.0f &&
                     (p.emitter[1] == 0 || strategy == LW_EMITTER_SAMPLING_NONE || S.bounces == 0);
  if (emits) S.radiance = S.radiance + mul(S.beta, mul(albedo, albedo)) * emittance;
  S.bounces += 1;

  const V3 origin = pos + n * eps;
  if (lit && dot(n, sun) > -sin_r && p.sun_dir[1] > -sin_r) {
    const V3 l = sample_sun(p, S.rng);
    const float cos_l = dot(n, l);
    if (cos_l > 0.0f && l.y >= 0.0f) {
      const V3 tr = shadow(s, p, lut, origin, l, INFINITY, n_traces);
      S.radiance = S.radiance + mul(mul(mul(S.beta, albedo), sun_color), tr) * cos_l;
    }
  }
  if (!emits && strategy != LW_EMITTER_SAMPLING_NONE) {
    S.radiance = S.radiance + mul(mul(S.beta, albedo), sample_emitters(s, p, lut, pos, origin, n, S.rng, n_traces));
  }

  S.beta = mul(S.beta, albedo);
  if (S.bounces >= 3) {
    const float q_rr = f_clamp(max_elem(S.beta), 0.05f, 0.95f);
    if (S.rng.next() > q_rr) return false;
    S.beta = S.beta / q_rr;
  }
  S.d = cosine_hemisphere(n, S.rng);
  S.o = origin;
  S.tmin = 0.0f;
  S.specular = false;
  return true;
}

// integrator.rs:539-709: the radiance of sample `sample` of pixel (px, py); the camera ray
// comes from `c` (the Params themselves, or one view of a batch).
template <class Cam>
__device__ __forceinline__ V3 render_sample(const DevScene &s, const lw_params &p, const Cam &c,
                                            const float *lut, uint32_t px, uint32_t py,
                                            uint32_t sample) {
  SampleState S;
  sample_begin(S, c, px, py, sample);
  while (sample_event(S, s, p, lut)) {
  }
  return S.radiance;
}

__device__ __forceinline__ V3 render_sample(const DevScene &s, const lw_params &p, const float *lut,
                                            uint32_t px, uint32_t py, uint32_t sample) {
  return render_sample(s, p, p, lut, px, py, sample);
}

}  // namespace lw
