/*
 * lumenwright-hip host-side frame preparation (C ABI).
 *
 * The reference prepares each HQ frame on the host before the backend call: the sun and
 * the Preetham sky (crates/lumenwright/src/render/sky.rs), the camera basis
 * (crates/lumenwright/src/scene/camera.rs), the kernel's Params (request.rs:1153-1225), the
 * sky cache (sky.rs:593-611 over integrator.rs:133-148) and the scene buffers
 * (scene/voxel.rs). These are the C++ counterparts, so a caller without the reference's
 * Rust crate (this repository's tests and bench) can build a frame. f64 code follows the
 * reference with the same libm calls; f32 code follows its evaluation order with no FMA
 * contraction.
 *
 * The synthetic worlds stand in for Minecraft saves, which need the game's client jar: a
 * value-noise terrain laid out exactly as build_hq lays out a loaded world
 * (scene/voxel.rs:111-491, emitter_list :542-607).
 */
#ifndef LW_HOST_H
#define LW_HOST_H

#include "lw/abi.h"

#ifdef __cplusplus
extern "C" {
#endif

/* sky.rs:28-32: vanilla celestial angle (0 = noon) for a time of day in ticks. */
float lw_celestial_angle(int64_t ticks);
/* sky.rs:35-38: unit vector towards the sun for a celestial angle. */
void lw_sun_direction_from_angle(float angle, float out[3]);
/* sky.rs:41-44: sun direction from altitude/azimuth in degrees. */
void lw_sun_direction(float altitude_deg, float azimuth_deg, float out[3]);
/* sky.rs:625-629 */
float lw_twilight(float altitude);

/* HqSky (sky.rs:549-557), built by hq_sky (sky.rs:636-644). */
typedef struct lw_hq_sky {
  float sun_dir[3];
  float sun_color[3];
  float sky_light;
} lw_hq_sky;
void lw_hq_sky_new(const float sun_dir[3], float sun_intensity, float sky_intensity,
                   lw_hq_sky *out);

/* Preetham (sky.rs:656-664), built by Preetham::new (sky.rs:666-768) at the given
 * turbidity (the reference fixes it at 2.5, sky.rs:667). */
typedef struct lw_preetham {
  float sun[3];
  float zenith[3];
  float perez[5][3];
} lw_preetham;
void lw_preetham_new(const float sun_dir[3], double turbidity, lw_preetham *out);

/* sky.rs:574-584 */
void lw_sun_texture_axes(const float sw[3], float su[3], float sv[3]);

/* sky.rs:593-611: the sky cache for params (resolution params.misc_f[2]); writes
 * (n + 1)^2 * 3 floats. Returns the count written, or a negative error if cap is short. */
long lw_sky_cache(const lw_params *params, float *out, size_t cap);

/* The 256 gamma-decoded byte values: gamma_to_linear(k / 255) (math.rs:213-216,
 * scene.rs:299-306), with libm's expf. */
void lw_gamma_lut(float out[256]);

/* The f32 e^x the device kernels evaluate (Beer-Lambert transmittance, math.rs:298-302),
 * computed here on the host for testing against libm's expf. */
void lw_expf_tabulated(const float *x, size_t n, float *out);

/* Camera::basis (camera.rs:20-37) for Minecraft yaw/pitch in degrees. */
void lw_camera_basis(float yaw_deg, float pitch_deg, float fwd[3], float right[3], float up[3]);

/* HqOptions (request.rs:180-210) as hq_params reads them. */
typedef struct lw_hq_options {
  uint32_t spp;
  uint32_t max_bounces;
  uint32_t seed;
  uint32_t emitter_sampling; /* LW_EMITTER_SAMPLING_* */
  uint32_t prevent_normal_emitter;
} lw_hq_options;

/* CameraRays (render/mod.rs:172-183). */
typedef struct lw_camera_rays {
  float forward[3];
  float right[3];
  float up[3];
  float tan_half_fov;
  float aspect;
} lw_camera_rays;

/* The camera rays of a frame (request.rs:643-652): Camera::basis, tan(fov_y / 2) with
 * fov_y in degrees, and width / height. */
void lw_camera_rays_new(float yaw_deg, float pitch_deg, float fov_y_deg, uint32_t width,
                        uint32_t height, lw_camera_rays *out);

/* hq_params (request.rs:1153-1225), with the Preetham sky at `turbidity`. */
void lw_hq_params(const lw_hq_options *o, float emitter_intensity, uint32_t width, uint32_t height,
                  const lw_scene_info *scene, const float eye[3], const lw_camera_rays *rays,
                  const lw_hq_sky *sky, double turbidity, lw_params *out);

/* ---- post-processing (crates/lumenwright/src/render/post.rs) ---- */

/* Tone-mapping operators (post.rs:3-21). */
#define LW_TONEMAP_GAMMA 0u
#define LW_TONEMAP_ACES 1u
#define LW_TONEMAP_FILMIC 2u
#define LW_TONEMAP_FILMIC_MID_GREY 3u

/* The Hable curve an operator uses (post.rs:37-138: Chunky's FILMIC_WORLDS preset, or it
 * brightened by mid_grey_gain(0.18) for FilmicMidGrey): a, b, c, d, e, f, w, then
 * white_scale() in out[7]. */
void lw_hable_curve(uint32_t op, float out[8]);

/* The byte thresholds of the display encoding (post.rs:141-144,165): out[k - 1] is the
 * smallest f32 v with (clamp(v, 0, 1)^(1/2.2) * 255 + 0.5) as u32 >= k, with libm's powf
 * (Rust std's f32::powf on Linux). The byte of v is then the number of thresholds <= v. */
void lw_encode_thresholds(float out[255]);

/* output.rs:19-30: writes hdr (w*h*3 floats, top row first) as a Portable Float Map,
 * rows bottom to top. Returns 0, or -1 with lw_last_error set. */
int lw_write_pfm(const char *path, const float *hdr, uint32_t w, uint32_t h);

/* ---- synthetic worlds ---- */

typedef struct lw_world_spec {
  uint32_t chunks_x, chunks_z; /* footprint in 16-block chunks */
  uint32_t sections;           /* height in 16-block sections */
  uint64_t seed;               /* terrain, block types, emitters, plants */
  uint64_t texture_seed;       /* atlas texels and emittances */
  uint32_t cube_types;         /* opaque full-cube block types */
  uint32_t cutout_types;       /* alpha-tested cross plants */
  int32_t water_level;         /* non-solid cells below this height hold water; < 0: none */
  float emitter_fraction;      /* of solid cells with a non-solid neighbour */
  float cutout_fraction;       /* of columns whose surface is above the water */
  uint32_t emitter_blocks;     /* this many more emitters, each the surface block of a seeded
                                  column with z >= half the depth (the half the default camera
                                  faces), emittance 1 on every face */
} lw_world_spec;

/* The HQ scene of a synthetic world (an owned HqScene). */
typedef struct lw_world {
  uint32_t *grid;
  size_t grid_len;
  lw_model *models;
  size_t models_len;
  lw_quad *quads;
  size_t quads_len;
  uint32_t *atlas;
  size_t atlas_len;
  lw_emitter *emitters;
  size_t emitters_len;
  uint32_t *emitter_grid;
  size_t emitter_grid_len;
  lw_scene_info info;
  /* per quad (quads_len each): the block type it belongs to (cube types 0..cube_types-1,
   * then the cutout types; LW_NO_GROUP for the water surface), and the emitter it belongs to
   * (the emitter's index in `emitters`, LW_NO_GROUP for none) */
  uint32_t *quad_type;
  uint32_t *quad_emitter;
  /* what the builder made (voxel.rs HqStats) */
  uint64_t bricks, uniform_bricks, buried_bricks, solid_cells, water_cells, cutout_cells;
  /* the interned volume the layout was made from, as lw_hip_scene_build takes it (see
   * lw_voxel_scene): voxels_len = 4096 per brick, brick_occludes one per brick,
   * model_flags one per model */
  uint16_t *voxels;
  size_t voxels_len;
  uint8_t *brick_occludes;
  uint8_t *model_flags;
} lw_world;

/* The terrain height h(x, z) = floor(64 + 24 V(x, z)) of a spec, V the 4-octave value noise. */
int32_t lw_world_height(const lw_world_spec *spec, int32_t x, int32_t z);
int lw_world_build(const lw_world_spec *spec, lw_world **out);
void lw_world_free(lw_world *w);

/* The directions of a w x h equirectangular map's texels (map_direction, sky.rs:856-860)
 * as tables: az[2i], az[2i+1] = cos, sin of column i's azimuth (i + 1/2) / w * tau and
 * el[2j], el[2j+1] = cos, sin of row j's height (j + 1/2) / h * pi - pi/2, with libm's cosf
 * and sinf; texel (i, j) looks along (el_cos * az_cos, el_sin, el_cos * az_sin). */
void lw_map_trig(uint32_t w, uint32_t h, float *az, float *el);

/* ---- the game's sky (render/sky.rs): the client's quick-mode sky drawn into maps ---- */

#define LW_LENS_MAP_W 2048u /* sky.rs:776 */
#define LW_LENS_MAP_H 1024u
#define LW_LIGHT_MAP_W 360u /* sky.rs:781 */
#define LW_LIGHT_MAP_H 180u
#define LW_MAX_SKY_LIGHT 50.0f   /* sky.rs:786 */
#define LW_DEFAULT_BRIGHTNESS 0.5f /* sky.rs:771 */
#define LW_STAR_BINS_AZ 128u     /* sky.rs:67-70 */
#define LW_STAR_BINS_EL 64u
#define LW_STAR_WORDS 9u

/* SkyDescription (sky.rs:225-253). */
typedef struct lw_sky_description {
  float sky[3];          /* visual/sky_color, sRGB 0..1 */
  float fog[3];          /* the fog colour, sRGB 0..1 */
  float sky_fog_end;
  float render_fog[2];   /* start, end */
  float env_fog[2];
  uint32_t sky_disc;
  float sunrise[4];      /* sRGB and alpha */
  float sun_angle;       /* degrees */
  uint32_t has_sun;
  float sun_dir[3];
  float sun_rect[4];     /* atlas x, y, w, h in texels */
  uint32_t has_moon;
  float moon_dir[3];
  float moon_rect[4];
  float star_angle;      /* degrees */
  float star_brightness;
} lw_sky_description;

/* QuickSky (sky.rs:256-283), the same layout. */
typedef struct lw_quick_sky {
  float sky_color[4];
  float fog_color[4];
  float fog_dist[4];
  float sunrise[4];
  float flags[4];
  float sun[4][4];
  float moon[4][4];
  float star_rot[3][4];
  float fan[18][4];
  float cam_forward[4];
  float cam_right[4];
  float cam_up[4];
  float cam_misc[4];
} lw_quick_sky;

/* QuickSky::new (sky.rs:320-373) for a camera and an output size. */
void lw_quick_sky_new(const lw_camera_rays *cam, uint32_t width, uint32_t height,
                      const lw_sky_description *d, lw_quick_sky *out);

/* celestial_direction (sky.rs:64-66): the unit vector towards a body at a celestial angle. */
void lw_celestial_direction(float angle_deg, float out[3]);

/* star_table (sky.rs:152-222): the client's stars binned by direction; returns the word
 * count and writes them to out when cap holds them (out may be null). */
long lw_star_table(uint32_t *out, size_t cap);

/* The sine and cosine (libm) of each sky-disc sector's edges: th0 = -pi + k pi/4 and
 * th0 + pi/4 for k = 0..7 (QuickSky::disc, sky.rs:408-410). */
void lw_disc_sector_trig(float s0[8], float c0[8], float s1[8], float c1[8]);

/* lightmap_sky (sky.rs:788-815): the client's lightmap colour for a cell open to the sky. */
void lw_lightmap_sky(float factor, const float colour[3], float brightness, float out[3]);

/* A procedural dusk sky as the game look's two equirectangular maps (the lens's, then the
 * light's; layout of integrator.rs:190-243): writes lw*lh*3 + gw*gh*3 floats. */
long lw_dusk_sky_maps(const float sun_dir[3], uint32_t lw, uint32_t lh, uint32_t gw, uint32_t gh,
                      float *out, size_t cap);

#ifdef __cplusplus
}
#endif

#endif /* LW_HOST_H */
