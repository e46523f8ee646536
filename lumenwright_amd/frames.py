"""Host-side frame preparation (include/lw/host.h) and the synthetic benchmark frames.

The reference builds a frame from a Minecraft save and the game's client jar
(crates/lumenwright/src/request.rs:751-950). Without them, the frames here are built from
seeded synthetic worlds with the same buffer layout (scene/voxel.rs) and the reference's
own sun, sky, camera and Params logic, restated in C++ (lumenwright_amd/csrc/lw_host.cpp).

Configurations (the workloads the benchmark and the parity tests run):
  0  8 x 8 chunks x 384, 32 cube types, 320 x 180, 16 spp, 4 bounces, Chunky's Preetham day
     sky at tick 4000 with the sun (sky cache), no emitters.
  1  24 x 24 chunks x 384, 56 cube + 8 cutout types, water below y = 62, ~1% emitters
     (intensity 137.9628, sampling ONE), 1280 x 720, 256 spp, 8 bounces, the game look's dusk
     at tick 12800 (two sky maps, no sun light).
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass, replace
from typing import Optional

import numpy as np

from . import _native
from .abi import (EMITTER, EMITTER_SAMPLING_NONE, EMITTER_SAMPLING_ONE, MODEL, PARAMS, QUAD,
                  QUICK_VERTEX, HqFrame, HqScene, VoxelVolume)

f3 = ctypes.c_float * 3


class HqSkyC(ctypes.Structure):
    _fields_ = [("sun_dir", ctypes.c_float * 3), ("sun_color", ctypes.c_float * 3),
                ("sky_light", ctypes.c_float)]


class PreethamC(ctypes.Structure):
    _fields_ = [("sun", ctypes.c_float * 3), ("zenith", ctypes.c_float * 3),
                ("perez", (ctypes.c_float * 3) * 5)]


class HqOptionsC(ctypes.Structure):
    _fields_ = [("spp", ctypes.c_uint32), ("max_bounces", ctypes.c_uint32),
                ("seed", ctypes.c_uint32), ("emitter_sampling", ctypes.c_uint32),
                ("prevent_normal_emitter", ctypes.c_uint32)]


class SceneInfoC(ctypes.Structure):
    _fields_ = [("dims", ctypes.c_uint32 * 3), ("offsets", ctypes.c_uint32 * 5),
                ("atlas_size", ctypes.c_uint32 * 2), ("quads_len", ctypes.c_uint32),
                ("models_len", ctypes.c_uint32), ("emitter_count", ctypes.c_uint32),
                ("see_through", ctypes.c_uint32 * 2), ("has_sun_sprite", ctypes.c_uint32),
                ("sun_sprite", ctypes.c_uint32 * 4)]


class CameraRaysC(ctypes.Structure):
    _fields_ = [("forward", ctypes.c_float * 3), ("right", ctypes.c_float * 3),
                ("up", ctypes.c_float * 3), ("tan_half_fov", ctypes.c_float),
                ("aspect", ctypes.c_float)]


class SkyDescriptionC(ctypes.Structure):
    """SkyDescription (sky.rs:225-253), lw_sky_description."""
    _fields_ = [("sky", ctypes.c_float * 3), ("fog", ctypes.c_float * 3),
                ("sky_fog_end", ctypes.c_float), ("render_fog", ctypes.c_float * 2),
                ("env_fog", ctypes.c_float * 2), ("sky_disc", ctypes.c_uint32),
                ("sunrise", ctypes.c_float * 4), ("sun_angle", ctypes.c_float),
                ("has_sun", ctypes.c_uint32), ("sun_dir", ctypes.c_float * 3),
                ("sun_rect", ctypes.c_float * 4), ("has_moon", ctypes.c_uint32),
                ("moon_dir", ctypes.c_float * 3), ("moon_rect", ctypes.c_float * 4),
                ("star_angle", ctypes.c_float), ("star_brightness", ctypes.c_float)]


class QuickSkyC(ctypes.Structure):
    """QuickSky (sky.rs:256-283), lw_quick_sky."""
    _fields_ = [("sky_color", ctypes.c_float * 4), ("fog_color", ctypes.c_float * 4),
                ("fog_dist", ctypes.c_float * 4), ("sunrise", ctypes.c_float * 4),
                ("flags", ctypes.c_float * 4), ("sun", (ctypes.c_float * 4) * 4),
                ("moon", (ctypes.c_float * 4) * 4), ("star_rot", (ctypes.c_float * 4) * 3),
                ("fan", (ctypes.c_float * 4) * 18), ("cam_forward", ctypes.c_float * 4),
                ("cam_right", ctypes.c_float * 4), ("cam_up", ctypes.c_float * 4),
                ("cam_misc", ctypes.c_float * 4)]


class WorldSpecC(ctypes.Structure):
    _fields_ = [("chunks_x", ctypes.c_uint32), ("chunks_z", ctypes.c_uint32),
                ("sections", ctypes.c_uint32), ("seed", ctypes.c_uint64),
                ("texture_seed", ctypes.c_uint64), ("cube_types", ctypes.c_uint32),
                ("cutout_types", ctypes.c_uint32), ("water_level", ctypes.c_int32),
                ("emitter_fraction", ctypes.c_float), ("cutout_fraction", ctypes.c_float),
                ("emitter_blocks", ctypes.c_uint32)]


class WorldC(ctypes.Structure):
    _fields_ = [("grid", ctypes.c_void_p), ("grid_len", ctypes.c_size_t),
                ("models", ctypes.c_void_p), ("models_len", ctypes.c_size_t),
                ("quads", ctypes.c_void_p), ("quads_len", ctypes.c_size_t),
                ("atlas", ctypes.c_void_p), ("atlas_len", ctypes.c_size_t),
                ("emitters", ctypes.c_void_p), ("emitters_len", ctypes.c_size_t),
                ("emitter_grid", ctypes.c_void_p), ("emitter_grid_len", ctypes.c_size_t),
                ("info", SceneInfoC),
                ("quad_type", ctypes.c_void_p), ("quad_emitter", ctypes.c_void_p),
                ("bricks", ctypes.c_uint64), ("uniform_bricks", ctypes.c_uint64),
                ("buried_bricks", ctypes.c_uint64), ("solid_cells", ctypes.c_uint64),
                ("water_cells", ctypes.c_uint64), ("cutout_cells", ctypes.c_uint64),
                ("voxels", ctypes.c_void_p), ("voxels_len", ctypes.c_size_t),
                ("brick_occludes", ctypes.c_void_p), ("model_flags", ctypes.c_void_p)]


def _v3(a) -> "ctypes.Array":
    return f3(*[float(x) for x in a])


# ---- sky.rs / camera.rs / request.rs restated ----

def celestial_angle(ticks: int) -> float:
    return _native.lib().lw_celestial_angle(ticks)


def sun_direction_from_angle(angle: float) -> np.ndarray:
    out = f3()
    _native.lib().lw_sun_direction_from_angle(angle, out)
    return np.array(out, dtype="<f4")


def sun_direction(altitude_deg: float, azimuth_deg: float) -> np.ndarray:
    out = f3()
    _native.lib().lw_sun_direction(altitude_deg, azimuth_deg, out)
    return np.array(out, dtype="<f4")


def twilight(altitude: float) -> float:
    return _native.lib().lw_twilight(altitude)


def hq_sky(sun_dir, sun_intensity: float = 1.25, sky_intensity: float = 1.0) -> HqSkyC:
    out = HqSkyC()
    _native.lib().lw_hq_sky_new(_v3(sun_dir), sun_intensity, sky_intensity, ctypes.byref(out))
    return out


def preetham_new(sun_dir, turbidity: float = 2.5) -> PreethamC:
    out = PreethamC()
    _native.lib().lw_preetham_new(_v3(sun_dir), turbidity, ctypes.byref(out))
    return out


def sun_texture_axes(sw):
    su, sv = f3(), f3()
    _native.lib().lw_sun_texture_axes(_v3(sw), su, sv)
    return np.array(su, dtype="<f4"), np.array(sv, dtype="<f4")


def sky_cache(params: np.ndarray) -> np.ndarray:
    n = int(params["misc_f"][2])
    if n == 0:
        return np.zeros(0, dtype="<f4")
    out = np.zeros((n + 1) * (n + 1) * 3, dtype="<f4")
    p = np.ascontiguousarray(params).reshape(())
    r = _native.lib().lw_sky_cache(ctypes.c_void_p(p.ctypes.data), ctypes.c_void_p(out.ctypes.data),
                                   out.size)
    assert r == out.size
    return out


def celestial_direction(angle_deg: float) -> np.ndarray:
    out = f3()
    _native.lib().lw_celestial_direction(angle_deg, out)
    return np.array(out, dtype="<f4")


def camera_rays(yaw: float, pitch: float, fov_y: float, width: int, height: int) -> "CameraRaysC":
    rays = CameraRaysC()
    _native.lib().lw_camera_rays_new(yaw, pitch, fov_y, width, height, ctypes.byref(rays))
    return rays


def sky_description(sky, fog, sky_fog_end, render_fog, env_fog, sky_disc, sunrise, sun_angle,
                    sun=None, moon=None, star_angle=0.0, star_brightness=0.0) -> SkyDescriptionC:
    """A SkyDescription; `sun` and `moon` are (direction, atlas rect) or None."""
    d = SkyDescriptionC()
    d.sky[:], d.fog[:] = list(sky), list(fog)
    d.sky_fog_end = sky_fog_end
    d.render_fog[:], d.env_fog[:] = list(render_fog), list(env_fog)
    d.sky_disc = int(sky_disc)
    d.sunrise[:] = list(sunrise)
    d.sun_angle = sun_angle
    if sun is not None:
        d.has_sun = 1
        d.sun_dir[:], d.sun_rect[:] = [float(x) for x in sun[0]], list(sun[1])
    if moon is not None:
        d.has_moon = 1
        d.moon_dir[:], d.moon_rect[:] = [float(x) for x in moon[0]], list(moon[1])
    d.star_angle, d.star_brightness = star_angle, star_brightness
    return d


def quick_sky_new(cam: "CameraRaysC", width: int, height: int, desc: SkyDescriptionC) -> QuickSkyC:
    """QuickSky::new (sky.rs:320-373)."""
    q = QuickSkyC()
    _native.lib().lw_quick_sky_new(ctypes.byref(cam), width, height, ctypes.byref(desc), ctypes.byref(q))
    return q


def star_table() -> np.ndarray:
    """star_table (sky.rs:152-222) as uint32 words."""
    L = _native.lib()
    n = L.lw_star_table(None, 0)
    out = np.zeros(n, dtype="<u4")
    L.lw_star_table(ctypes.c_void_p(out.ctypes.data), n)
    return out


def lightmap_sky(factor: float, colour, brightness: float = 0.5) -> np.ndarray:
    out = f3()
    _native.lib().lw_lightmap_sky(factor, _v3(colour), brightness, out)
    return np.array(out, dtype="<f4")


def map_trig(w: int, h: int):
    """lw_map_trig: (cos, sin) per column azimuth and per row height of a w x h map."""
    az = np.zeros((w, 2), dtype="<f4")
    el = np.zeros((h, 2), dtype="<f4")
    _native.lib().lw_map_trig(w, h, ctypes.c_void_p(az.ctypes.data), ctypes.c_void_p(el.ctypes.data))
    return az, el


def gamma_lut() -> np.ndarray:
    out = np.zeros(256, dtype="<f4")
    _native.lib().lw_gamma_lut(ctypes.c_void_p(out.ctypes.data))
    return out


def camera_basis(yaw: float, pitch: float):
    f, r, u = f3(), f3(), f3()
    _native.lib().lw_camera_basis(yaw, pitch, f, r, u)
    return (np.array(f, dtype="<f4"), np.array(r, dtype="<f4"), np.array(u, dtype="<f4"))


def scene_info(scene: HqScene) -> SceneInfoC:
    info = SceneInfoC()
    info.dims[:] = list(scene.dims)
    info.offsets[:] = list(scene.offsets)
    info.atlas_size[:] = list(scene.atlas_size)
    info.quads_len = len(scene.quads)
    info.models_len = len(scene.models)
    info.emitter_count = scene.emitter_count
    info.see_through[:] = list(scene.see_through)
    if scene.sun_sprite is not None:
        info.has_sun_sprite = 1
        info.sun_sprite[:] = list(scene.sun_sprite)
    return info


def hq_params(scene: HqScene, *, width: int, height: int, eye, yaw: float, pitch: float,
              fov_y: float, sky: HqSkyC, max_bounces: int, seed: int = 0,
              emitter_intensity: float = 0.0, emitter_sampling: int = EMITTER_SAMPLING_ONE,
              prevent_normal_emitter: bool = False, turbidity: float = 2.5) -> np.ndarray:
    """hq_params (request.rs:1153-1225) for a camera at `eye` (grid-relative) with Minecraft
    yaw/pitch and a vertical fov in degrees (request.rs:643-652)."""
    rays = CameraRaysC()
    _native.lib().lw_camera_rays_new(yaw, pitch, fov_y, width, height, ctypes.byref(rays))
    opts = HqOptionsC(0, max_bounces, seed, emitter_sampling, int(prevent_normal_emitter))
    out = np.zeros((), dtype=PARAMS)
    _native.lib().lw_hq_params(ctypes.byref(opts), emitter_intensity, width, height,
                               ctypes.byref(scene_info(scene)), _v3(eye), ctypes.byref(rays),
                               ctypes.byref(sky), turbidity, ctypes.c_void_p(out.ctypes.data))
    return out


def dusk_sky_maps(sun_dir, lens=(2048, 1024), light=(360, 180)) -> np.ndarray:
    n = (lens[0] * lens[1] + light[0] * light[1]) * 3
    out = np.zeros(n, dtype="<f4")
    r = _native.lib().lw_dusk_sky_maps(_v3(sun_dir), lens[0], lens[1], light[0], light[1],
                                       ctypes.c_void_p(out.ctypes.data), n)
    assert r == n
    return out


# ---- post-processing host helpers (render/post.rs, lumenwright-cli output.rs) ----

TONEMAP_GAMMA, TONEMAP_ACES, TONEMAP_FILMIC, TONEMAP_FILMIC_MID_GREY = 0, 1, 2, 3


def hable_curve(op: int) -> np.ndarray:
    """The operator's Hable curve (a, b, c, d, e, f, w) and its white scale (post.rs:37-138)."""
    out = np.zeros(8, dtype="<f4")
    _native.lib().lw_hable_curve(op, ctypes.c_void_p(out.ctypes.data))
    return out


def encode_thresholds() -> np.ndarray:
    """The 255 byte thresholds of the display encoding (libm powf)."""
    out = np.zeros(255, dtype="<f4")
    _native.lib().lw_encode_thresholds(ctypes.c_void_p(out.ctypes.data))
    return out


def write_pfm(path: str, hdr: np.ndarray, w: int, h: int) -> None:
    """output.rs:19-30: a Portable Float Map, rows bottom to top."""
    a = np.ascontiguousarray(hdr, dtype="<f4").reshape(-1)
    assert a.size == w * h * 3
    if _native.lib().lw_write_pfm(str(path).encode(), ctypes.c_void_p(a.ctypes.data), w, h) != 0:
        raise OSError(f"writing {path}")


# ---- synthetic worlds ----

@dataclass(frozen=True)
class WorldSpec:
    chunks_x: int = 8
    chunks_z: int = 8
    sections: int = 24
    seed: int = 1
    texture_seed: int = 2
    cube_types: int = 32
    cutout_types: int = 0
    water_level: int = -1
    emitter_fraction: float = 0.0
    cutout_fraction: float = 0.0
    emitter_blocks: int = 0

    def c(self) -> WorldSpecC:
        return WorldSpecC(self.chunks_x, self.chunks_z, self.sections, self.seed,
                          self.texture_seed, self.cube_types, self.cutout_types, self.water_level,
                          self.emitter_fraction, self.cutout_fraction, self.emitter_blocks)


def world_height(spec: WorldSpec, x: int, z: int) -> int:
    return _native.lib().lw_world_height(ctypes.byref(spec.c()), x, z)


def _copy(ptr, n, dtype):
    if n == 0:
        return np.zeros(0, dtype=dtype)
    nbytes = n * np.dtype(dtype).itemsize
    buf = (ctypes.c_uint8 * nbytes).from_address(ptr)
    return np.frombuffer(bytes(buf), dtype=dtype).copy()


def build_world(spec: WorldSpec, volume: bool = False) -> HqScene:
    """The synthetic world's HQ scene (lw_world_build). With volume=True it also carries
    `.volume`, the interned voxel volume the layout was made from (a VoxelVolume, what
    lw_hip_scene_build takes)."""
    L = _native.lib()
    h = ctypes.c_void_p()
    rc = L.lw_world_build(ctypes.byref(spec.c()), ctypes.byref(h))
    if rc != 0:
        raise RuntimeError("world build failed")
    try:
        w = WorldC.from_address(h.value)
        info = w.info
        scene = HqScene(
            grid=_copy(w.grid, w.grid_len, "<u4"),
            models=_copy(w.models, w.models_len, MODEL),
            quads=_copy(w.quads, w.quads_len, QUAD),
            atlas=_copy(w.atlas, w.atlas_len, "<u4"),
            atlas_size=(info.atlas_size[0], info.atlas_size[1]),
            dims=tuple(info.dims),
            offsets=tuple(info.offsets),
            see_through=tuple(info.see_through),
            emitters=_copy(w.emitters, w.emitters_len, EMITTER),
            emitter_count=info.emitter_count,
            emitter_grid=_copy(w.emitter_grid, w.emitter_grid_len, "<u4"),
        )
        scene.stats = {k: int(getattr(w, k)) for k in ("bricks", "uniform_bricks", "buried_bricks",
                                                        "solid_cells", "water_cells",
                                                        "cutout_cells")}
        # Per quad: its block type and emitter (NO_GROUP for none), for differentiable
        # rendering's parameter groups.
        scene.quad_type = _copy(w.quad_type, w.quads_len, "<u4")
        scene.quad_emitter = _copy(w.quad_emitter, w.quads_len, "<u4")
        if volume:
            nb = int(info.dims[0]) * int(info.dims[1]) * int(info.dims[2])
            scene.volume = VoxelVolume(
                dims=tuple(info.dims), voxels=_copy(w.voxels, w.voxels_len, "<u2"),
                models=scene.models, model_flags=_copy(w.model_flags, w.models_len, np.uint8),
                quads=scene.quads, atlas=scene.atlas, atlas_size=scene.atlas_size,
                brick_occludes=_copy(w.brick_occludes, nb, np.uint8))
        return scene
    finally:
        L.lw_world_free(h)


GAME_EMITTER_INTENSITY = 137.9628  # request.rs:138
DEFAULT_EMITTER_INTENSITY = 13.0   # request.rs:126


@dataclass(frozen=True)
class FrameConfig:
    """One benchmark workload: a world, a camera and the HQ options."""
    name: str
    world: WorldSpec
    width: int
    height: int
    spp: int
    max_bounces: int
    sky: str                 # "chunky" (Preetham cache, sun light) or "game-dusk" (two maps)
    ticks: int
    emitter_intensity: float
    fov_y: float = 70.0
    pitch: float = 30.0
    yaw: float = 0.0
    camera_height: float = 20.0
    seed: int = 0
    turbidity: float = 2.5
    # The sun from altitude and azimuth in radians (sky.rs:41-44) instead of the time of day.
    sun_altitude: Optional[float] = None
    sun_azimuth: Optional[float] = None


CONFIGS = {
    0: FrameConfig("cfg0: 8x8 chunks, 320x180, 16 spp, 4 bounces, Preetham day sky",
                   WorldSpec(8, 8, 24, cube_types=32), 320, 180, 16, 4, "chunky", 4000, 0.0),
    1: FrameConfig("cfg1: 24x24 chunks with water and emitters, 1280x720, 256 spp, 8 bounces, "
                   "game dusk sky",
                   WorldSpec(24, 24, 24, cube_types=56, cutout_types=8, water_level=62,
                             emitter_fraction=0.01, cutout_fraction=0.05),
                   1280, 720, 256, 8, "game-dusk", 12800, GAME_EMITTER_INTENSITY),
    # The differentiable render: the cfg0 world with 16 emitter blocks (emittance 1 times the
    # default emitter intensity, request.rs:126), its sun (tick 4000: altitude 67.955 deg,
    # azimuth 90 deg) given by angles so that they can be parameters.
    3: FrameConfig("cfg3: differentiable render, 8x8 chunks with 16 emitter blocks, 256x144, "
                   "64 spp, 4 bounces, Preetham day sky",
                   WorldSpec(8, 8, 24, cube_types=32, emitter_blocks=16), 256, 144, 64, 4,
                   "chunky", 4000, DEFAULT_EMITTER_INTENSITY,
                   sun_altitude=1.186048, sun_azimuth=1.5707963),
}


def config(n: int, **overrides) -> FrameConfig:
    """CONFIGS[n] with fields replaced; `world_*` keys replace fields of the world spec."""
    c = CONFIGS[n]
    wo = {k[6:]: v for k, v in overrides.items() if k.startswith("world_")}
    fo = {k: v for k, v in overrides.items() if not k.startswith("world_")}
    if wo:
        fo["world"] = replace(c.world, **wo)
    return replace(c, **fo)


_world_cache = {}


def make_frame(cfg: FrameConfig, scene: Optional[HqScene] = None) -> HqFrame:
    """The HqFrame of a configuration: its world, camera, Params and sky buffer."""
    if scene is None:
        scene = _world_cache.get(cfg.world)
        if scene is None:
            scene = build_world(cfg.world)
            _world_cache.clear()
            _world_cache[cfg.world] = scene
    gx, gz = cfg.world.chunks_x * 16, cfg.world.chunks_z * 16
    cx, cz = gx // 2, gz // 2
    eye = [cx + 0.5, world_height(cfg.world, cx, cz) + cfg.camera_height, cz + 0.5]
    if cfg.sun_altitude is not None and cfg.sun_azimuth is not None:
        import math
        sun = sun_direction(math.degrees(cfg.sun_altitude), math.degrees(cfg.sun_azimuth))
    else:
        sun = sun_direction_from_angle(celestial_angle(cfg.ticks))
    sky = hq_sky(sun, 1.25, 1.0)
    params = hq_params(scene, width=cfg.width, height=cfg.height, eye=eye, yaw=cfg.yaw,
                       pitch=cfg.pitch, fov_y=cfg.fov_y, sky=sky, max_bounces=cfg.max_bounces,
                       seed=cfg.seed, emitter_intensity=cfg.emitter_intensity,
                       turbidity=cfg.turbidity)
    if cfg.sky == "chunky":
        sky_buf = sky_cache(params)
    elif cfg.sky == "game-dusk":
        # The game look (request.rs:848-924 at day weight 0): the sun casts no light and is
        # not drawn by Chunky's sprite, the maps replace the cache.
        params["sun_color"][:3] = 0.0
        params["sun_sprite"][:] = 0.0
        params["sky_map"] = [2048.0, 1024.0, 360.0, 180.0]
        params["misc_f"][2] = 0.0
        sky_buf = dusk_sky_maps(sun)
    else:
        raise ValueError(cfg.sky)
    return HqFrame(scene=scene, params=params, spp=cfg.spp, sky=sky_buf)


def batch_views(cfg: FrameConfig, scene: HqScene, n_views: int = 64, seed: int = 3,
                sizes=((160, 90), (320, 180), (640, 360), (1280, 720)),
                spps=(16, 64, 256, 1024), max_bounces: int = 8):
    """The batched workload over a world: `n_views` seeded cameras, each at a point 10 to 40
    blocks above the terrain looking at a seeded surface point, with a resolution and spp
    drawn uniformly from `sizes` and `spps`; per-view seed = view index. Returns
    [(params, spp)] sharing the frame's sky; every Params field but the camera, the size
    and the seed is the base frame's."""
    import math
    base = make_frame(replace(cfg, max_bounces=max_bounces), scene)
    rng = np.random.default_rng(seed)
    gx, gz = cfg.world.chunks_x * 16, cfg.world.chunks_z * 16
    views = []
    for v in range(n_views):
        w, h = sizes[int(rng.integers(len(sizes)))]
        spp = int(spps[int(rng.integers(len(spps)))])
        ex, ez = rng.uniform(16, gx - 16), rng.uniform(16, gz - 16)
        ey = world_height(cfg.world, int(ex), int(ez)) + rng.uniform(10, 40)
        tx, tz = rng.uniform(16, gx - 16), rng.uniform(16, gz - 16)
        ty = world_height(cfg.world, int(tx), int(tz)) + 0.5
        dx, dy, dz = tx - ex, ty - ey, tz - ez
        norm = math.sqrt(dx * dx + dy * dy + dz * dz)
        dx, dy, dz = dx / norm, dy / norm, dz / norm
        pitch = math.degrees(math.asin(-dy))          # camera.rs:40-45 (look_at)
        yaw = math.degrees(math.atan2(-dx, dz))
        rays = CameraRaysC()
        _native.lib().lw_camera_rays_new(yaw, pitch, cfg.fov_y, w, h, ctypes.byref(rays))
        p = base.params.copy()
        p["cam_pos"] = [ex, ey, ez, rays.tan_half_fov]
        p["cam_fwd"] = list(rays.forward) + [rays.aspect]
        p["cam_right"] = list(rays.right) + [0.0]
        p["cam_up"] = list(rays.up) + [0.0]
        p["frame"][0], p["frame"][1] = w, h
        p["misc_u"][0] = v
        views.append((p, spp))
    return base, views


# ---- quick mode: synthetic frames ----

QUICK_TILE = 16  # texels per block texture side


def quick_view_proj(yaw: float, pitch: float, fov_y: float, aspect: float, near: float,
                    far: float) -> np.ndarray:
    """Camera::projection(aspect, near, far) * Camera::view(0) (scene/camera.rs:47-70), glam's
    Mat4 as 16 floats, column-major: the camera-relative view-projection of a quick frame
    (request.rs:697-698)."""
    f, r, u = camera_basis(yaw, pitch)
    view = np.zeros((4, 4), "<f4")  # view[c] = column c
    view[0] = [r[0], u[0], -f[0], 0.0]
    view[1] = [r[1], u[1], -f[1], 0.0]
    view[2] = [r[2], u[2], -f[2], 0.0]
    view[3] = [0.0, 0.0, 0.0, 1.0]
    h = np.float32(1.0) / np.float32(np.tan(np.float32(np.radians(np.float32(fov_y))) * np.float32(0.5)))
    w = h / np.float32(aspect)
    rr = np.float32(far) / (np.float32(near) - np.float32(far))
    proj = np.zeros((4, 4), "<f4")
    proj[0, 0], proj[1, 1], proj[2, 2], proj[2, 3], proj[3, 2] = w, h, rr, -1.0, rr * np.float32(near)
    out = np.zeros((4, 4), "<f4")
    for c in range(4):
        v = view[c]
        acc = proj[0] * v[0]
        acc = acc + proj[1] * v[1]
        acc = acc + proj[2] * v[2]
        out[c] = acc + proj[3] * v[3]
    return out.reshape(16)


def _box_down(level: np.ndarray) -> np.ndarray:
    """A 2 x 2 box filter of packed RGBA8 texels (h, w) -> (h / 2, w / 2)."""
    ch = np.stack([(level >> s) & 0xFF for s in (0, 8, 16, 24)], -1).astype(np.float32)
    m = (ch[0::2, 0::2] + ch[1::2, 0::2] + ch[0::2, 1::2] + ch[1::2, 1::2]) * 0.25
    m = np.minimum(np.rint(m), 255).astype(np.uint32)
    return (m[..., 0] | (m[..., 1] << 8) | (m[..., 2] << 16) | (m[..., 3] << 24)).astype("<u4")


@dataclass
class QuickAtlas:
    """A block atlas with its mip chain (assets/atlas.rs's layout): 16 x 16 block textures in
    the mipmapped rows at the top (cube, cutout and water tiles), then a region read nearest
    (text glyphs and the sun sprite)."""
    pixels: np.ndarray
    size: tuple
    mips: np.ndarray
    levels: int
    mip_rows: int
    cube: list
    cutout: list
    water: tuple
    text: tuple   # (x, y, w, h) texels
    sun: tuple    # (x, y, w, h) texels


def quick_atlas(cube_types: int = 56, cutout_types: int = 8, seed: int = 2, levels: int = 4) -> QuickAtlas:
    rng = np.random.default_rng(seed)
    T, per_row = QUICK_TILE, 16
    n_tiles = cube_types + cutout_types + 1
    block_rows = (n_tiles + per_row - 1) // per_row
    mip_rows = block_rows * T
    W, H = per_row * T, mip_rows + 64
    img = np.zeros((H, W), dtype="<u4")

    def rgba(r, g, b, a):
        return (r.astype(np.uint32) | (g.astype(np.uint32) << 8) | (b.astype(np.uint32) << 16)
                | (np.asarray(a, np.uint32) << 24))

    tiles = []
    for k in range(n_tiles):
        x, y = (k % per_row) * T, (k // per_row) * T
        base = rng.integers(40, 216, 3)
        noise = rng.integers(-30, 31, (T, T, 3))
        c = np.clip(base[None, None, :] + noise, 0, 255)
        a = np.full((T, T), 255)
        if cube_types <= k < cube_types + cutout_types:
            a = np.where(rng.random((T, T)) < 0.4, 0, 255)  # leaves: holes for the alpha test
        elif k == n_tiles - 1:
            c = np.clip(np.array([50, 90, 200])[None, None, :] + noise // 3, 0, 255)
            a = np.full((T, T), 180)
        img[y:y + T, x:x + T] = rgba(c[..., 0], c[..., 1], c[..., 2], a)
        tiles.append((x, y))
    # text glyphs (white ink, a dark see-through backdrop) and the sun
    tx, ty, tw, th = 0, mip_rows, 128, 32
    ink = rng.random((th, tw)) < 0.3
    img[ty:ty + th, tx:tx + tw] = np.where(ink, 0xFFFFFFFF, 0x40000000).astype("<u4")
    sx, sy = 128, mip_rows
    yy, xx = np.mgrid[0:T, 0:T]
    disc = (xx - 7.5) ** 2 + (yy - 7.5) ** 2 < 36
    img[sy:sy + T, sx:sx + T] = np.where(disc, 0xFF80F0FF, 0xFF000000).astype("<u4")
    mips, lv = [], img
    for _ in range(levels):
        lv = _box_down(lv)
        mips.append(lv.reshape(-1))
    return QuickAtlas(pixels=img.reshape(-1), size=(W, H), mips=np.concatenate(mips) if mips else
                      np.zeros(0, "<u4"), levels=levels, mip_rows=mip_rows,
                      cube=tiles[:cube_types], cutout=tiles[cube_types:cube_types + cutout_types],
                      water=tiles[-1], text=(tx, ty, tw, th), sun=(sx, sy, T, T))


# Per face: outward normal, the corner offset of its base and its two edges (t1 x t2 = normal,
# so the corners base, +t1, +t1+t2, +t2 wind counter-clockwise seen from outside).
_FACES = {
    "+y": ((0, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 0), 1.0),
    "-y": ((0, -1, 0), (0, 0, 0), (1, 0, 0), (0, 0, 1), 0.5),
    "+x": ((1, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), 0.6),
    "-x": ((-1, 0, 0), (0, 0, 0), (0, 0, 1), (0, 1, 0), 0.6),
    "+z": ((0, 0, 1), (0, 0, 1), (1, 0, 0), (0, 1, 0), 0.8),
    "-z": ((0, 0, -1), (0, 0, 0), (0, 1, 0), (1, 0, 0), 0.8),
}


def _hash3(x, y, z):
    h = (np.asarray(x, np.int64) * 73856093) ^ (np.asarray(y, np.int64) * 19349663) ^ (np.asarray(z, np.int64) * 83492791)
    return (h & 0x7FFFFFFF).astype(np.int64)


class _QuadSink:
    def __init__(self, atlas: QuickAtlas, eye):
        self.atlas, self.eye = atlas, np.asarray(eye, np.float64)
        self.parts = []

    def faces(self, face: str, blocks: np.ndarray, tiles: np.ndarray, alpha=255, y_off=0.0, rect=None):
        """One quad per block (n, 3 integer corners) on `face`, textured with atlas tiles (n, 2)
        (or one rect (x, y, w, h) for all)."""
        if len(blocks) == 0:
            return
        _, base, t1, t2, shade = _FACES[face]
        p0 = blocks.astype(np.float64) + np.array(base, np.float64) + np.array([0.0, y_off, 0.0])
        t1, t2 = np.array(t1, np.float64), np.array(t2, np.float64)
        corners = np.stack([p0, p0 + t1, p0 + t1 + t2, p0 + t2], 1) - self.eye  # (n, 4, 3)
        W, H = self.atlas.size
        if rect is None:
            u0, v0 = tiles[:, 0] / W, tiles[:, 1] / H
            u1, v1 = (tiles[:, 0] + QUICK_TILE) / W, (tiles[:, 1] + QUICK_TILE) / H
        else:
            u0 = np.full(len(blocks), rect[0] / W)
            v0 = np.full(len(blocks), rect[1] / H)
            u1 = np.full(len(blocks), (rect[0] + rect[2]) / W)
            v1 = np.full(len(blocks), (rect[1] + rect[3]) / H)
        uv = np.stack([np.stack([u0, v1], -1), np.stack([u1, v1], -1), np.stack([u1, v0], -1),
                       np.stack([u0, v0], -1)], 1)
        ao = np.array([1.0, 0.8, 0.9, 0.7])[(_hash3(blocks[:, 0], blocks[:, 1], blocks[:, 2])[:, None]
                                             + np.arange(4)[None, :]) % 4]
        light = np.clip(np.rint(255.0 * shade * 0.92 * ao), 0, 255)
        v = np.zeros((len(blocks), 4), dtype=QUICK_VERTEX)
        v["pos"] = corners.astype("<f4")
        v["uv"] = uv.astype("<f4")
        v["color"][..., 0] = light
        v["color"][..., 1] = light
        v["color"][..., 2] = light
        v["color"][..., 3] = alpha
        self.parts.append(v)

    def take(self):
        """The quads so far as (vertices (4n), centres (n, 3)); clears the sink."""
        if not self.parts:
            return np.zeros(0, QUICK_VERTEX), np.zeros((0, 3))
        v = np.concatenate(self.parts)
        self.parts = []
        return v.reshape(-1), v["pos"].astype(np.float64).mean(axis=1)


def _quad_indices(n: int, first: int) -> np.ndarray:
    q = first + 4 * np.arange(n, dtype=np.int64)[:, None]
    return (q + np.array([0, 1, 2, 0, 2, 3])[None, :]).astype("<u4").reshape(-1)


def quick_frame(cfg: "FrameConfig", draw_distance: float = 128.0, width: Optional[int] = None,
                height: Optional[int] = None, displays: int = 3):
    """A quick-mode frame of the configuration's world and camera (request.rs:691-747 with the
    QuickOptions defaults: draw distance 128, 4 mip levels, no supersampling), the mesh made
    from the synthetic world's height field: every exposed face of the terrain and of leaf
    blocks (opaque and cutout), water surfaces below the world's water level (translucent, back
    to front) and text displays before the camera (see-through, far to near). Chunks outside
    the view frustum or the draw distance are left out, as the reference's mesh builder leaves
    them out (scene/mesh.rs:329). Returns (QuickFrame, atlas)."""
    from .abi import QuickFrame, QuickMesh
    spec = cfg.world
    W, H = width or cfg.width, height or cfg.height
    gx, gz = spec.chunks_x * 16, spec.chunks_z * 16
    cx, cz = gx // 2, gz // 2
    eye = np.array([cx + 0.5, world_height(spec, cx, cz) + cfg.camera_height, cz + 0.5])
    near, far = 0.05, draw_distance * 2.0 + 512.0
    vp = quick_view_proj(cfg.yaw, cfg.pitch, cfg.fov_y, W / H, near, far)
    atlas = quick_atlas(max(spec.cube_types, 1), max(spec.cutout_types, 1), spec.texture_seed)
    R = int(draw_distance)
    x0, x1 = max(cx - R, 0), min(cx + R, gx - 1)
    z0, z1 = max(cz - R, 0), min(cz + R, gz - 1)
    xs, zs = np.arange(x0, x1 + 1), np.arange(z0, z1 + 1)
    L = _native.lib()
    sc = spec.c()
    hm = np.array([[L.lw_world_height(ctypes.byref(sc), int(x), int(z)) for z in zs] for x in xs], np.int64)
    X, Z = np.meshgrid(xs, zs, indexing="ij")
    # chunk culling: the draw distance, then the frustum (all 8 corners outside one plane)
    m = vp.reshape(4, 4).astype(np.float64)  # rows are columns
    keep = np.zeros_like(hm, dtype=bool)
    for bx in range(x0 // 16, x1 // 16 + 1):
        for bz in range(z0 // 16, z1 // 16 + 1):
            sel = (X // 16 == bx) & (Z // 16 == bz)
            if not sel.any():
                continue
            dx = max(bx * 16 - eye[0], 0, eye[0] - (bx * 16 + 16))
            dz = max(bz * 16 - eye[2], 0, eye[2] - (bz * 16 + 16))
            if np.hypot(dx, dz) > draw_distance:
                continue
            y_lo, y_hi = hm[sel].min() - 1, hm[sel].max() + 3
            c = np.array([[bx * 16 + i * 16, y, bz * 16 + k * 16] for i in (0, 1) for k in (0, 1)
                          for y in (y_lo, y_hi)], np.float64) - eye
            clip = c @ m[:3] + m[3]
            if (np.all(clip[:, 0] > clip[:, 3]) or np.all(clip[:, 0] < -clip[:, 3]) or
                    np.all(clip[:, 1] > clip[:, 3]) or np.all(clip[:, 1] < -clip[:, 3]) or
                    np.all(clip[:, 2] < 0)):
                continue
            keep |= sel
    n_cube = len(atlas.cube)
    cube = np.array(atlas.cube, np.int64)
    sink = _QuadSink(atlas, eye)
    kx, kz, kh = X[keep], Z[keep], hm[keep]
    top = np.stack([kx, kh, kz], -1)
    sink.faces("+y", top, cube[_hash3(kx, kh, kz) % n_cube])
    pad = np.pad(hm, 1, mode="edge")
    for face, (ox, oz) in (("+x", (1, 0)), ("-x", (-1, 0)), ("+z", (0, 1)), ("-z", (0, -1))):
        hn = pad[1 + ox:1 + ox + hm.shape[0], 1 + oz:1 + oz + hm.shape[1]][keep]
        depth = np.maximum(kh - hn, 0)
        rep = np.repeat(np.arange(kh.size), depth)
        y = kh[rep] - (np.arange(rep.size) - np.repeat(np.cumsum(depth) - depth, depth))
        blk = np.stack([kx[rep], y, kz[rep]], -1)
        sink.faces(face, blk, cube[_hash3(blk[:, 0], blk[:, 1] // 4, blk[:, 2]) % n_cube])
    # leaves: a two-block cutout column on a few dry columns
    wl = spec.water_level
    leafy = (_hash3(kx, 7, kz) % 23 == 0) & (kh >= wl)
    cut = np.array(atlas.cutout, np.int64)
    for dy in (1, 2):
        blk = np.stack([kx[leafy], kh[leafy] + dy, kz[leafy]], -1)
        t = cut[_hash3(blk[:, 0], blk[:, 1], blk[:, 2]) % len(cut)]
        for face in _FACES:
            sink.faces(face, blk, t)
    opaque_v, _ = sink.take()
    # water surfaces, back to front
    wet = kh < wl
    blk = np.stack([kx[wet], np.full(wet.sum(), wl), kz[wet]], -1)
    sink.faces("+y", blk, np.tile(np.array(atlas.water, np.int64), (len(blk), 1)), y_off=-0.125)
    water_v, centres = sink.take()
    order = np.argsort(-np.einsum("ij,ij->i", centres, centres), kind="stable")
    water_v = water_v.reshape(-1, 4)[order].reshape(-1)
    # text displays facing the camera, far to near
    f, r, u = (np.asarray(a, np.float64) for a in camera_basis(cfg.yaw, cfg.pitch))
    disp = []
    for k in range(displays):
        centre = f * (10.0 - 2.5 * k) + r * (1.5 * (k - 1)) + u * 0.6 * k
        p0 = centre - r * 2.0 - u * 0.5
        corners = np.stack([p0, p0 + 4.0 * r, p0 + 4.0 * r + u, p0 + u])
        tx, ty, tw, th = atlas.text
        Wa, Ha = atlas.size
        v = np.zeros(4, QUICK_VERTEX)
        v["pos"] = corners.astype("<f4")
        v["uv"] = np.array([[tx, ty + th], [tx + tw, ty + th], [tx + tw, ty], [tx, ty]], np.float64) / [Wa, Ha]
        v["color"] = [255, 255, 255, 255]
        disp.append(v)
    see_v = np.concatenate(disp) if disp else np.zeros(0, QUICK_VERTEX)
    verts = np.concatenate([opaque_v, water_v, see_v])
    n_o, n_w, n_s = opaque_v.size // 4, water_v.size // 4, see_v.size // 4
    mesh = QuickMesh(vertices=verts, opaque=_quad_indices(n_o, 0),
                     translucent=_quad_indices(n_w, 4 * n_o),
                     see_through=_quad_indices(n_s, 4 * (n_o + n_w)))
    sun = celestial_direction(90.29)
    desc = sky_description(
        sky=(0.18, 0.24, 0.37), fog=(0.62, 0.40, 0.33), sky_fog_end=draw_distance,
        render_fog=(draw_distance - min(max(draw_distance / 10.0, 4.0), 64.0), draw_distance),
        env_fog=(0.0, 1024.0), sky_disc=True, sunrise=(0.98, 0.57, 0.25, 0.62), sun_angle=90.29,
        sun=(sun, atlas.sun), star_angle=90.29, star_brightness=0.38)
    sky = quick_sky_new(camera_rays(cfg.yaw, cfg.pitch, cfg.fov_y, W, H), W, H, desc)
    frame = QuickFrame(mesh=mesh, atlas_pixels=atlas.pixels, atlas_size=atlas.size,
                       atlas_mips=atlas.mips, mip_levels=atlas.levels, mip_rows=atlas.mip_rows,
                       stars=star_table(), width=W, height=H, view_proj=vp, sky=sky)
    return frame, atlas
