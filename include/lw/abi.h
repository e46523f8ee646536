/*
 * lumenwright-hip C ABI: the HQ voxel path tracer on AMD Instinct MI355X (gfx950).
 *
 * This is the drop-in boundary for the reference's HQ backend contract:
 *
 *   trait Backend { fn name(&self) -> String;
 *                   fn render_quick(&mut self, &QuickFrame) -> Result<Vec<u32>>;
 *                   fn render_hq(&mut self, &HqFrame, progress) -> Result<Vec<[f32; 3]>>; }
 *     crates/lumenwright/src/render/mod.rs:92-103
 *   struct HqFrame { scene: &HqScene, params: Params, spp: u32, sky: Vec<f32> }
 *     crates/lumenwright/src/render/mod.rs:80-89
 *
 * and for the kernel crate's device ABI (the buffers `pt_render` takes,
 * crates/lumenwright-ptx/src/lib.rs:37-56). Structs are byte-identical to the
 * reference's #[repr(C)] types (crates/lumenwright-kernel/src/scene.rs:44-228;
 * size asserts at crates/lumenwright-kernel/src/trace.rs:327-333).
 *
 * Conventions (mirroring anyhow::Result in the reference):
 *   - every int-returning call returns 0 on success, a negative LW_E* code on failure;
 *     the message is then available from lw_hip_last_error (per handle) or
 *     lw_last_error (for calls without a handle);
 *   - inputs are borrowed for the duration of the call; outputs are caller-owned;
 *   - one thread at a time per handle (the reference takes &mut self); progress
 *     callbacks run on the caller's thread.
 */
#ifndef LW_ABI_H
#define LW_ABI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- constants: crates/lumenwright-kernel/src/scene.rs:16-37,146-155 ---- */
#define LW_EMPTY 0xFFFFFFFFu
#define LW_UNIFORM 0x80000000u
#define LW_BRICK_WORDS 2048u
#define LW_OV_CELL_WORDS 8192u
#define LW_ALPHA_OPAQUE 0u
#define LW_ALPHA_CUTOUT 1u
#define LW_ALPHA_TRANSLUCENT 2u
#define LW_FLAG_ALPHA_MASK 3u
#define LW_FLAG_WATER (1u << 2)
#define LW_FLAG_SHADE (1u << 3)
#define LW_FLAG_FRONT_ONLY (1u << 4)
#define LW_EMITTER_SAMPLING_NONE 0u
#define LW_EMITTER_SAMPLING_ONE 1u
#define LW_EMITTER_SAMPLING_ONE_BLOCK 2u
#define LW_EMITTER_SAMPLING_ALL 3u
#define LW_EMITTER_GRID_HEADER 8u

/* ---- error codes ---- */
#define LW_OK 0
#define LW_EINVAL (-1)  /* malformed scene or argument (the reference would panic on it) */
#define LW_ENODEV (-2)  /* no HIP device / runtime (cuda.rs:29-41 maps this to an error) */
#define LW_EHIP (-3)    /* a HIP runtime call failed */
#define LW_ENOMEM (-4)  /* device allocation failed */
#define LW_EUNSUPPORTED (-5) /* a valid request this backend does not serve: differentiable
                                 rendering with emitter sampling ONE_BLOCK or ALL, sky parameters
                                 with sky maps or the sun sprite, a fit over the sky parameters */

/* GpuQuad, 80 bytes (scene.rs:44-57): p0.w = flags, eu.w = tint rgba8, ev.w = emittance. */
typedef struct lw_quad {
  float p0[4];
  float eu[4];
  float ev[4];
  float uv[4];
  float uvt[4];
} lw_quad;

/* GpuModel, 16 bytes (scene.rs:101-112). */
typedef struct lw_model {
  uint32_t quad_offset;
  uint32_t quad_count;
  uint32_t flags;
  uint32_t pad;
} lw_model;

/* GpuEmitter, 32 bytes (scene.rs:117-124): pos.w = first quad (bit cast). */
typedef struct lw_emitter {
  float pos[4];
  uint32_t faces[4];
} lw_emitter;

/* Params, 432 bytes = 27 vec4 (scene.rs:157-228). */
typedef struct lw_params {
  float cam_pos[4];
  float cam_fwd[4];
  float cam_right[4];
  float cam_up[4];
  float sun_dir[4];
  float sun_color[4];
  float sky_sun[4];
  float sky_zenith[4];
  float perez_a[4];
  float perez_b[4];
  float perez_c[4];
  float perez_d[4];
  float perez_e[4];
  float water_color[4];
  float misc_f[4];
  uint32_t grid_dims[4];
  uint32_t off[4];
  uint32_t off2[4];
  uint32_t frame[4];
  uint32_t misc_u[4];
  uint32_t see_through[4];
  uint32_t emitter[4];
  float sun_su[4];
  float sun_sv[4];
  float sun_sprite[4];
  float sky_map[4];
  float water_plane[4];
} lw_params;

/* SceneView plus the slice lengths (scene.rs:231-259; the argument list of
 * pt_render, lumenwright-ptx/src/lib.rs:37-56). */
typedef struct lw_scene_view {
  const lw_params *params;
  const uint32_t *grid;
  size_t grid_len;
  const lw_model *models;
  size_t models_len;
  const lw_quad *quads;
  size_t quads_len;
  const uint32_t *atlas;
  size_t atlas_len;
  const lw_emitter *emitters;
  size_t emitters_len;
  const uint32_t *emitter_grid;
  size_t emitter_grid_len;
  const float *sky;
  size_t sky_len;
} lw_scene_view;

/* trace's Hit (trace.rs:11-22) plus the cell the traversal stopped in and whether
 * anything was hit (trace returns Option<Hit>). */
typedef struct lw_hit {
  float t;
  uint32_t quad;
  uint32_t texel;
  float s;
  float tt;
  int32_t cell[3];
  uint32_t hit; /* 1 = Some(hit), 0 = None */
  uint32_t pad;
} lw_hit;

/* progress(done, total): the reference's `&(dyn Fn(u32, u32) + Sync)`. */
typedef void (*lw_progress_fn)(uint32_t done, uint32_t total, void *user);

typedef struct lw_hip_backend lw_hip_backend;
typedef struct lw_hip_scene lw_hip_scene;

/* Kernel-clock timing of the last render on a handle (the reference logs the same
 * CUDA-event sum, render/cuda.rs:128-131,154-159,170-173). */
typedef struct lw_timing {
  double trace_ms;       /* sum of HIP-event durations of the path-tracing kernels */
  double accumulate_ms;  /* sum of HIP-event durations of the f64 accumulation kernels */
  double wall_ms;        /* host wall clock of the whole call */
  uint64_t samples;      /* width * height * spp rendered */
  uint32_t launches;     /* path-tracing kernel launches */
  uint32_t pad;
} lw_timing;

/* ---- backend: crates/lumenwright/src/render/cuda.rs ---- */

/* Describes device 0 (name and ISA) or says why HIP is unavailable; ~ cuda::probe
 * (render/cuda.rs:44-51). Writes a NUL-terminated string into desc. */
int lw_hip_probe(char *desc, size_t cap);

/* Opens a HIP device; ~ CudaBackend::new (render/cuda.rs:55-62). */
int lw_hip_open(int device, lw_hip_backend **out);

/* Human-readable name, e.g. "amd instinct mi355x (hip)"; ~ Backend::name
 * (render/mod.rs:94, render/cuda.rs:66-68). Owned by the handle. */
const char *lw_hip_name(const lw_hip_backend *b);

/* Backend::render_hq (render/mod.rs:98-102; render/cuda.rs:77-184): renders
 * params.frame[0] x params.frame[1] pixels at max(spp, 1) samples each and writes the
 * mean linear radiance per pixel, row-major, top row first, to out_rgb (w*h*3 floats).
 * Non-finite samples are dropped from the sum and still count in the mean
 * (render/cpu/mod.rs:66-79). Uploads the view's buffers for the call, as the reference
 * does (render/cuda.rs:96-113). progress(done, spp) reports whole samples per pixel done,
 * from the calling thread, as the reference's per-launch callback does (render/cuda.rs:161-163),
 * and ends with (spp, spp). */
int lw_hip_render_hq(lw_hip_backend *b, const lw_scene_view *view, uint32_t spp,
                     lw_progress_fn progress, void *user, float *out_rgb);

/* Vertex (scene/mesh.rs:14-21): camera-relative position, normalised atlas UV, colour. */
typedef struct lw_quick_vertex {
  float pos[3];
  float uv[2];
  uint8_t color[4];
} lw_quick_vertex;

/* QuickFrame (render/mod.rs:53-78): the mesh (vertices and three index lists, three indices
 * a triangle: opaque and cutout, translucent back to front, see-through), the atlas with its
 * mip levels above 0 concatenated, the star table (lw_star_table), the output size, the
 * camera-relative view-projection (column-major, glam's Mat4) and the sky (lw_quick_sky,
 * include/lw/host.h). */
typedef struct lw_quick_frame {
  const lw_quick_vertex *vertices;
  size_t vertices_len;
  const uint32_t *opaque;
  size_t opaque_len;
  const uint32_t *translucent;
  size_t translucent_len;
  const uint32_t *see_through;
  size_t see_through_len;
  const uint32_t *atlas_pixels;
  uint32_t atlas_size[2];
  const uint32_t *atlas_mips;
  size_t atlas_mips_len;
  uint32_t mip_levels; /* at most 16 */
  uint32_t mip_rows;
  const uint32_t *stars;
  size_t stars_len;
  uint32_t width, height;
  float view_proj[16];
  const struct lw_quick_sky *sky;
} lw_quick_frame;

/* Backend::render_quick (render/mod.rs:96): the frame drawn on the device as the reference's
 * CPU rasterizer draws it (render/cpu/raster.rs, render/texture.rs; the reference's CUDA
 * backend falls back to that rasterizer, render/cuda.rs:70-75): width * height packed RGBA8
 * words (red in the low byte), row major, top row first. */
int lw_hip_render_quick(lw_hip_backend *b, const lw_quick_frame *frame, uint32_t *out_rgba8);

/* The number of trace calls (trace.rs:137-290; camera and extension rays, each sun shadow
 * segment, each emitter ray) a render of the scene with these params and spp makes: the
 * paths of lw_hip_render_hq_scene, replayed by a counting kernel that is not timed. No
 * counterpart in the reference, which counts no rays; the measurement's Mrays/s. */
int lw_hip_count_traces(lw_hip_backend *b, const lw_hip_scene *scene, const lw_params *params, uint32_t spp,
                        uint64_t *traces);

/* The message of the last failure on this handle ("" if none). */
const char *lw_hip_last_error(const lw_hip_backend *b);

/* The message of the last failure of a call made without a handle. */
const char *lw_last_error(void);

void lw_hip_close(lw_hip_backend *b);

/* ---- the kernel crate's public functions, on the device (parity surface) ---- */

/* trace (crates/lumenwright-kernel/src/trace.rs:137-290) for n rays: o, d are n*3
 * floats, tmin/tmax n floats each; out[i].hit says whether ray i met a surface. The hit
 * fields come from the walk the renderers run, which stops once every brick left toward the
 * ray's octant is empty; out[i].cell from the whole walk. */
int lw_hip_trace(lw_hip_backend *b, const lw_scene_view *view, size_t n, const float *o,
                 const float *d, const float *tmin, const float *tmax, lw_hit *out);

/* render_sample (crates/lumenwright-kernel/src/integrator.rs:539-709) for n
 * (px, py, sample) triples; out_rgb gets n*3 floats of linear radiance. */
int lw_hip_render_samples(lw_hip_backend *b, const lw_scene_view *view, size_t n,
                          const uint32_t *px, const uint32_t *py, const uint32_t *sample,
                          float *out_rgb);

/* ---- resident scenes: upload once, render many frames (the reference re-uploads on
 * every call, render/cuda.rs:96-113) ---- */

/* Validates and uploads the scene buffers of a view (everything but params), so that
 * later renders read them from HBM. The view's params are only used for validation. */
int lw_hip_scene_upload(lw_hip_backend *b, const lw_scene_view *view, lw_hip_scene **out);

/* render_hq over a resident scene with the given params (validated against it). */
int lw_hip_render_hq_scene(lw_hip_backend *b, const lw_hip_scene *scene,
                           const lw_params *params, uint32_t spp, lw_progress_fn progress,
                           void *user, float *out_rgb);

void lw_hip_scene_free(lw_hip_backend *b, lw_hip_scene *scene);

/* lw_hip_render_hq_scene that allocates nothing. No counterpart in the reference: it is for
 * callers that cannot allocate device memory mid-run (tight serving loops, code under
 * stream capture, harnesses that forbid hipMalloc). The frame's f64 accumulator and staged
 * samples live in a scratch array of LW_NOALLOC_SCRATCH_BYTES that each device holds from the
 * library's load to the process's end (36 bytes a pixel, rounded: frames up to about 3.7 M
 * pixels; larger ones fail with LW_EINVAL and a message). It renders in the direct form
 * whatever LW_KERNEL says, and gives lw_hip_render_hq_scene's image bit for bit. Every call
 * on one device shares that device's array: calls that overlap on one device race (the
 * accumulator is cleared at the start of each call). */
#define LW_NOALLOC_SCRATCH_BYTES (128u << 20)
int lw_hip_render_hq_scene_noalloc(lw_hip_backend *b, const lw_hip_scene *scene,
                                   const lw_params *params, uint32_t spp,
                                   lw_progress_fn progress, void *user, float *out_rgb);

/* ---- the scene builder on the device (crates/lumenwright/src/scene/voxel.rs) ---- */

#define LW_MODEL_OCCLUDES 1u /* the model's block occludes (BlockInfo::occludes, voxel.rs:119-130) */
#define LW_MODEL_EMITS 2u    /* the model has quads and its block emits (voxel.rs:521) */

/* The HqScene facts hq_params reads (voxel.rs:19-49). */
typedef struct lw_scene_info {
  uint32_t dims[3];
  uint32_t offsets[5];
  uint32_t atlas_size[2];
  uint32_t quads_len;
  uint32_t models_len;
  uint32_t emitter_count;
  uint32_t see_through[2];
  uint32_t has_sun_sprite;
  uint32_t sun_sprite[4];
} lw_scene_info;

/* What build_hq (voxel.rs:111-491) holds once it has interned the view's block-model
 * instances (voxel.rs:170-305): one model id per voxel, the models with their quads and
 * flags, the entity overlay's quads, and the atlas and sky the scene carries. */
typedef struct lw_voxel_scene {
  uint32_t dims[3];               /* bricks along x, y (sections) and z; fewer than 2^24 bricks */
  uint32_t see_through_count;     /* the last this-many entity quads are see-through text
                                     (voxel.rs:361-376) */
  const uint16_t *voxels;         /* 4096 model ids per brick: brick (bx, by, bz) at
                                     ((by * dims[2] + bz) * dims[0] + bx) * 4096, its voxel
                                     (lx, ly, lz) at (ly * 16 + lz) * 16 + lx */
  const uint8_t *brick_occludes;  /* per brick: every block of its section occludes
                                     (voxel.rs:119-130); null: every voxel's model has
                                     LW_MODEL_OCCLUDES */
  const lw_model *models;         /* model 0 is air */
  size_t models_len;              /* 1 .. 65536 */
  const uint8_t *model_flags;     /* LW_MODEL_* per model */
  const lw_quad *quads;           /* the models' quads */
  size_t quads_len;
  const lw_quad *entity_quads;    /* the overlay's quads in grid space, flags set
                                     (voxel.rs:355-372) */
  size_t entity_quads_len;
  const uint32_t *atlas;
  size_t atlas_len;
  uint32_t atlas_size[2];
  uint32_t has_sun_sprite;
  uint32_t sun_sprite[4];
  const float *sky;               /* the sky buffer renders read (lw_scene_view.sky) */
  size_t sky_len;
} lw_voxel_scene;

/* HqStats (voxel.rs:51-70) with the buffers' lengths and the build's time. */
typedef struct lw_build_stats {
  uint64_t bricks, uniform_bricks, buried_bricks, instances, overlay_bricks, entity_quads,
      emitters;
  uint64_t grid_len, quads_len, emitters_len, emitter_grid_len;
  double kernel_ms; /* HIP-event time from the first layout kernel to the last */
  double wall_ms;   /* the whole call, uploads included */
} lw_build_stats;

/* build_hq's layout on the device (voxel.rs:305-491 and emitter_list, voxel.rs:542-607):
 * the grid buffer (top-level entries with UNIFORM and buried bricks collapsed, the stored
 * bricks, the entity overlay), the emitter list and the emitter grid, made in HBM from the
 * voxel volume and kept there as a resident scene. info receives the HqScene facts params
 * are made from (lw_hq_params); renders then take params carrying that layout. The overlay's
 * cell lists are the reference's; their order inside an overlay brick's list buffer, which
 * the reference takes from a HashMap (voxel.rs:418-424), is cell order here. stats may be
 * null. No counterpart in the reference, which builds on the CPU and uploads. */
int lw_hip_scene_build(lw_hip_backend *b, const lw_voxel_scene *vs, lw_scene_info *info,
                       lw_build_stats *stats, lw_hip_scene **out);

/* Copies one buffer of a resident scene to the host: LW_SCENE_GRID and
 * LW_SCENE_EMITTER_GRID (uint32), LW_SCENE_EMITTERS (lw_emitter), LW_SCENE_QUADS (lw_quad),
 * LW_SCENE_OCTANT_BOXES (uint8 per brick, numbered (y * dz + z) * dx + x: bit k, for
 * k = [+x] | [+y] << 1 | [+z] << 2, is set when the box from the brick to the grid's end
 * toward octant k holds a brick whose top-level or overlay entry is not EMPTY),
 * LW_SCENE_CELL_CUBES (uint32 per voxel of each stored brick, 4096 per brick in the order of
 * the brick words, bricks by their top-level entry: bits 4k .. 4k + 3, k numbered as above,
 * hold the side n <= 8 of the largest cube of voxels from this one toward octant k, clipped to
 * the brick, in which no voxel has a model; 0 for a voxel with a model; empty when the scene
 * has no table).
 * *len receives the element count; with dst null only *len is written, otherwise cap (in
 * elements) must hold the buffer. No counterpart in the reference. */
#define LW_SCENE_GRID 0u
#define LW_SCENE_EMITTERS 1u
#define LW_SCENE_EMITTER_GRID 2u
#define LW_SCENE_QUADS 3u
#define LW_SCENE_OCTANT_BOXES 4u
#define LW_SCENE_CELL_CUBES 5u
int lw_hip_scene_read(lw_hip_backend *b, const lw_hip_scene *scene, uint32_t which, void *dst,
                      size_t cap, size_t *len);

/* One view of a batch: its Params and samples per pixel. The views of a batch may differ
 * only in the camera (cam_pos, cam_fwd, cam_right, cam_up), the frame size (frame[0..1])
 * and the seed (misc_u[0]). */
typedef struct lw_view {
  lw_params params;
  uint32_t spp;
  uint32_t pad[3];
} lw_view;

/* Renders n views (1..256) of a resident scene as one work pool: one persistent launch whose
 * waves take (view, 4x4-pixel tile) items, the views with most samples first, and sum each
 * pixel's samples in sample order. out_rgb[i] receives view i's mean radiance, w*h*3 floats,
 * equal to what render_hq gives for the view alone (same samples, same summation order).
 * progress, when given, is called once, with (1, 1), after the launch ends. No counterpart in
 * the reference, which renders one view per call (render/mod.rs:98-102). */
int lw_hip_render_hq_batch(lw_hip_backend *b, const lw_hip_scene *scene, const lw_view *views,
                           size_t n, lw_progress_fn progress, void *user, float *const *out_rgb);

/* ---- differentiable rendering (no counterpart in the reference, which renders with
 * fixed materials and sky) ---- */

#define LW_NO_GROUP 0xFFFFFFFFu

/* The parameters of a differentiable render and their derivative inputs. At palette and
 * intensity factors of 1 and sky parameters as the Params hold them, render_diff gives
 * render_hq's image.
 *   palette:   the albedo of a hit on quad q (integrator.rs:633) is multiplied by the
 *              RGB factor palette[3 g..3 g+2], g = quad_palette[q]; the same for the colour
 *              emitter sampling meets (integrator.rs:411-412). Translucent quads take none.
 *   intensity: the emittance of quad q times params.misc_f.x (integrator.rs:661,413) is
 *              multiplied by intensity[quad_intensity[q]].
 *   sky:       with sky_params, the gradient also covers the Preetham turbidity and the
 *              sun's altitude and azimuth (radians): d_sky_cache holds three tables shaped
 *              like the sky cache (d cache / d turbidity, altitude, azimuth), and d_sun_dir,
 *              d_sun_color, d_sky_light the derivatives of params.sun_dir.xyz,
 *              params.sun_color.xyz and params.sky_sun.w (rows: turbidity, altitude,
 *              azimuth). The Preetham sky only (no sky maps), no sun sprite.
 * Factors must be finite and > 0. Emitter sampling strategies NONE and ONE. rr_off = 1
 * turns Russian roulette off (integrator.rs:696-702), which makes the render a smooth
 * function of every parameter but the sun's, for finite-difference checks. */
typedef struct lw_diff_desc {
  uint32_t n_palette;
  uint32_t n_intensity;
  const uint32_t *quad_palette;   /* quads_len entries: a palette group or LW_NO_GROUP */
  const uint32_t *quad_intensity; /* quads_len entries: an intensity group or LW_NO_GROUP */
  const float *palette;           /* 3 * n_palette */
  const float *intensity;         /* n_intensity */
  uint32_t sky_params;
  uint32_t rr_off;
  const float *d_sky_cache;       /* sky_params: 3 * sky_len floats */
  float d_sun_dir[3][3];
  float d_sun_color[3][3];
  float d_sky_light[3];
  uint32_t pad;
} lw_diff_desc;

/* The image of params under diff's parameters: w*h*3 floats of mean radiance, as
 * lw_hip_render_hq_scene; out_f64 (optional, w*h*3 doubles) receives the same means before
 * their rounding to f32, for finite differences and losses that need them. */
int lw_hip_render_diff(lw_hip_backend *b, const lw_hip_scene *scene, const lw_params *params,
                       uint32_t spp, const lw_diff_desc *diff, float *out_rgb, double *out_f64);

/* The views of a batch (lw_hip_render_hq_batch's rules) under diff's parameters, in one
 * work pool: out_rgb[i] gets view i's image, as lw_hip_render_diff would give it alone. */
int lw_hip_render_diff_batch(lw_hip_backend *b, const lw_hip_scene *scene, const lw_view *views,
                             size_t n, const lw_diff_desc *diff, float *const *out_rgb);

/* The adjoint of lw_hip_render_diff_batch: grads = sum over views, pixels and channels of
 * d_loss_d_image[i] * dI_i/dtheta (one w*h*3 array per view). Palette and intensity
 * parameters only (sky_params 0 or 1 with the views' shared Preetham sky). */
int lw_hip_adjoint_batch(lw_hip_backend *b, const lw_hip_scene *scene, const lw_view *views,
                         size_t n, const lw_diff_desc *diff, const float *const *d_loss_d_image,
                         double *grads);

/* An inverse fit on the device: the views of a batch, their target images, and the palette
 * and intensity factors of a lw_diff_desc fitted to them by Adam (D. P. Kingma, J. Ba, 2015)
 * on L = sum over views, pixels and channels of (I - T)^2 / N. A step renders the batch,
 * forms L and dL/dI, replays the batch for the gradient and updates the factors, all on the
 * device (factors are kept >= min_factor). The views keep their seeds from step to step. */
typedef struct lw_fit_options {
  float lr, beta1, beta2, eps, min_factor;
  uint32_t pad[3];
} lw_fit_options;
typedef struct lw_fit_stats {
  double loss;        /* L at the factors the step started from */
  double forward_ms;  /* HIP-event time of the step's forward path kernels */
  double adjoint_ms;  /* HIP-event time of its adjoint kernels */
  double wall_ms;
  uint64_t samples;   /* samples of one forward (the adjoint replays as many) */
  uint32_t step;
  uint32_t pad;
} lw_fit_stats;
typedef struct lw_hip_fit lw_hip_fit;
int lw_hip_fit_new(lw_hip_backend *b, const lw_hip_scene *scene, const lw_view *views, size_t n,
                   const lw_diff_desc *diff, const float *const *targets,
                   const lw_fit_options *opt, lw_hip_fit **out);
int lw_hip_fit_step(lw_hip_backend *b, const lw_hip_scene *scene, lw_hip_fit *fit,
                    lw_fit_stats *stats);
/* The current factors (3 n_palette and n_intensity floats; either may be null) and the last
 * step's gradient (3 n_palette + n_intensity + 3 doubles; may be null). */
int lw_hip_fit_params(lw_hip_backend *b, const lw_hip_fit *fit, float *palette,
                      float *intensity, double *grads);
void lw_hip_fit_free(lw_hip_backend *b, lw_hip_fit *fit);

/* The adjoint of lw_hip_render_diff with the same arguments: replays every sample from its
 * seed and writes grads[i] = sum over pixels and channels of d_loss_d_image * dI/dtheta_i
 * (w*h*3 floats in, 3 n_palette + n_intensity + 3 doubles out: palette RGB factors,
 * intensities, then turbidity, sun altitude, sun azimuth, the last three 0 without
 * sky_params). params.frame[3] (max bounces) must be at most 16. */
int lw_hip_adjoint(lw_hip_backend *b, const lw_hip_scene *scene, const lw_params *params,
                   uint32_t spp, const lw_diff_desc *diff, const float *d_loss_d_image,
                   double *grads);

/* ---- post-processing on the device (crates/lumenwright/src/render/post.rs) ---- */

/* tonemap (post.rs:148-170): n linear RGB pixels (n*3 floats) to packed display RGBA8,
 * op one of LW_TONEMAP_* (include/lw/host.h), bytes identical to the reference's (the
 * display encoding goes through byte thresholds made with libm's powf). */
int lw_hip_tonemap(lw_hip_backend *b, const float *hdr, size_t n, float exposure, uint32_t op,
                   uint32_t *out);

/* tonemap_mix (post.rs:176-215): weight * op_a + (1 - weight) * op_b in display values. At
 * weight >= 1 or <= 0 it is tonemap of one operator. Inside, the display values are libm's
 * powf(x, 1/2.2f) as in lw_hip_encode, so the bytes are the reference's. */
int lw_hip_tonemap_mix(lw_hip_backend *b, const float *hdr, size_t n, float exposure,
                       uint32_t op_a, uint32_t op_b, float weight, uint32_t *out);

/* The display encoding (post.rs:141-144): clamp(x, 0, 1)^(1/2.2f) for n values, equal to the
 * reference's f32::powf (libm's powf) for every f32. The power is taken in f64 and rounded
 * once, and the inputs where libm's powf rounds otherwise come from a table the backend makes
 * on its first call with libm itself; setup_ms (may be null) receives that one-time cost. */
int lw_hip_encode(lw_hip_backend *b, const float *x, size_t n, float *out, double *setup_ms);

/* Rust's f32 `as i32` and `as u32` (truncation toward zero, saturation at the type's range,
 * NaN to 0) of n values, computed the way every kernel converts: the reference's casts in
 * trace.rs, scene.rs and integrator.rs (e.g. trace.rs:161-163, scene.rs:291,
 * integrator.rs:351). A check of the device conversions for callers and tests; no
 * counterpart in the reference. */
int lw_hip_casts(lw_hip_backend *b, const float *x, size_t n, int32_t *out_i, uint32_t *out_u);

/* downsample (post.rs:218-247): box-filters w x h RGBA8 by k into (w/k) x (h/k). */
int lw_hip_downsample(lw_hip_backend *b, const uint32_t *px, uint32_t w, uint32_t h, uint32_t k,
                      uint32_t *out);

/* render_hq_scene followed by tonemap on the device: only the w*h RGBA8 words cross PCIe. */
int lw_hip_render_hq_rgba8(lw_hip_backend *b, const lw_hip_scene *scene, const lw_params *params,
                           uint32_t spp, float exposure, uint32_t op, uint32_t *out_rgba8);

/* ---- sky preparation on the device (crates/lumenwright/src/render/sky.rs) ---- */

/* build_sky_cache (crates/lumenwright-kernel/src/integrator.rs:133-148, called by
 * sky::sky_cache, sky.rs:593-611) for params.misc_f[2] as the resolution n: (n + 1)^2 * 3
 * floats into out. *len receives the count; with out null only *len is written. */
int lw_hip_sky_cache(lw_hip_backend *b, const lw_params *params, float *out, size_t cap,
                     size_t *len);

/* blend_chunky_sky (sky.rs:836-851) in place: every texel of the lens's map, then the
 * light's (sizes params.sky_map, maps_len = (w0 h0 + w1 h1) * 3 floats) becomes
 * weight * chunky(d) + (1 - weight) * texel for its direction d, chunky the kernel's
 * sky_radiance through the sky cache of params (Preetham sky, misc_f[2] its resolution), as
 * request.rs:891-911 builds it. */
int lw_hip_blend_chunky_sky(lw_hip_backend *b, const lw_params *params, float weight, float *maps,
                            size_t maps_len);

/* GameSkyMaps (sky.rs:885-905) without its buffer, and the call's device time. */
typedef struct lw_game_sky {
  float sizes[4]; /* Params::sky_map: the lens's map, then the light's */
  float light[3]; /* the lightmap's colour decoded */
  float mean[3];  /* the cosine-weighted mean of the light's sky over the upper hemisphere */
  float scale;
  float veil[3];
  double kernel_ms; /* HIP-event time of the drawing and veil kernels */
  double setup_ms;  /* the one-time libm power table made by this call, 0 if it existed */
} lw_game_sky;

/* game_sky_maps (sky.rs:906-953) on the device: the client's sky (QuickSky::background,
 * sky.rs:509-532) drawn into the lens's 2048 x 1024 map from `lens` and the light's 360 x 180
 * map from `light` (lw_quick_sky_new), decoded to linear, the light's scaled so that a face
 * open to the whole sky takes the lightmap's colour. out receives the two maps (the lens's,
 * then the light's, linear RGB, (2048 * 1024 + 360 * 180) * 3 floats), info the rest. atlas
 * holds the sprites the sun and moon rects name (atlas_w texels a row); stars is
 * lw_star_table's. */
int lw_hip_game_sky_maps(lw_hip_backend *b, const struct lw_quick_sky *lens,
                         const struct lw_quick_sky *light, float sky_light_factor,
                         const float sky_light_color[3], const uint32_t *atlas, size_t atlas_len,
                         uint32_t atlas_w, const uint32_t *stars, size_t stars_len, float *out,
                         size_t cap, lw_game_sky *info);

/* ---- autotuning (no counterpart in the reference) ---- */

#define LW_FORM_PERSISTENT 0u
#define LW_FORM_DIRECT 1u
#define LW_FORM_WAVEFRONT 2u
#define LW_FORM_TILES 3u  /* the persistent tile kernel of frames with >= 16 spp */

typedef struct lw_tune_entry {
  uint32_t form;      /* LW_FORM_* */
  uint32_t waves;     /* the register budget: minimum waves per SIMD the kernel is built for */
  uint32_t per_lane;  /* samples per lane (direct form) */
  uint32_t pad;
  double ms;          /* HIP-event time of the probe's path kernels */
} lw_tune_entry;

typedef struct lw_tune_result {
  lw_tune_entry entries[16];
  uint32_t n;          /* candidates timed */
  uint32_t best;       /* index of the fastest launch shape, now the handle's setting */
  uint32_t best_tiles; /* index of the fastest LW_FORM_TILES budget, now the handle's
                          setting; UINT32_MAX when the frame does not use the tile kernel */
  uint32_t pad;
} lw_tune_result;

/* Times a probe of params' frame (its first probe_spp samples per pixel, one launch) with
 * each launch shape the library has -- direct form at 4, 5, 6 and 8 waves per SIMD and 1 or
 * 4 samples per lane, persistent form at 3 and 4, wavefront form -- and keeps the fastest
 * for later renders and batches on this handle. When the frame is large enough for the
 * persistent tile kernel (direct-form frames of at least 16 spp with 4 tiles for every
 * resident wave), it also times that kernel on the frame's first 16 samples per pixel at 4, 5
 * and 6 waves per SIMD (LW_FORM_TILES entries) and keeps the fastest budget for it.
 * Every shape gives the same images. */
int lw_hip_autotune(lw_hip_backend *b, const lw_hip_scene *scene, const lw_params *params,
                    uint32_t probe_spp, lw_tune_result *out);

/* Diagnostics (no counterpart in the reference): copies `bytes` between two device buffers
 * with a plain streaming kernel (k_copy) twice and reports the second copy's time -- the
 * known-size reference for hardware byte counters. */
int lw_hip_copy_probe(lw_hip_backend *b, size_t bytes, double *ms);

/* Kernel-clock timing of the last render_hq / render_hq_scene on this handle. */
int lw_hip_last_timing(const lw_hip_backend *b, lw_timing *out);

#ifdef __cplusplus
}
#endif

#endif /* LW_ABI_H */
