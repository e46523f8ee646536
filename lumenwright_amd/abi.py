"""Python mirror of include/lw/abi.h: numpy dtypes of the scene ABI and the ctypes view.

The record layouts are byte-identical to the reference's ``#[repr(C)]`` types
(crates/lumenwright-kernel/src/scene.rs:44-228; sizes asserted at
crates/lumenwright-kernel/src/trace.rs:327-333).
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

# scene.rs:16-37,146-155
EMPTY = 0xFFFFFFFF
UNIFORM = 0x80000000
BRICK_WORDS = 2048
OV_CELL_WORDS = 8192
ALPHA_OPAQUE = 0
ALPHA_CUTOUT = 1
ALPHA_TRANSLUCENT = 2
FLAG_ALPHA_MASK = 3
FLAG_WATER = 1 << 2
FLAG_SHADE = 1 << 3
FLAG_FRONT_ONLY = 1 << 4
EMITTER_SAMPLING_NONE = 0
EMITTER_SAMPLING_ONE = 1
EMITTER_SAMPLING_ONE_BLOCK = 2
EMITTER_SAMPLING_ALL = 3
EMITTER_GRID_HEADER = 8

# GpuQuad (scene.rs:44-57), 80 bytes.
QUAD = np.dtype([("p0", "<f4", 4), ("eu", "<f4", 4), ("ev", "<f4", 4), ("uv", "<f4", 4),
                 ("uvt", "<f4", 4)])
# GpuModel (scene.rs:101-112), 16 bytes.
MODEL = np.dtype([("quad_offset", "<u4"), ("quad_count", "<u4"), ("flags", "<u4"), ("pad", "<u4")])
# GpuEmitter (scene.rs:117-124), 32 bytes.
EMITTER = np.dtype([("pos", "<f4", 4), ("faces", "<u4", 4)])

_PARAM_FIELDS = [
    ("cam_pos", "<f4"), ("cam_fwd", "<f4"), ("cam_right", "<f4"), ("cam_up", "<f4"),
    ("sun_dir", "<f4"), ("sun_color", "<f4"), ("sky_sun", "<f4"), ("sky_zenith", "<f4"),
    ("perez_a", "<f4"), ("perez_b", "<f4"), ("perez_c", "<f4"), ("perez_d", "<f4"),
    ("perez_e", "<f4"), ("water_color", "<f4"), ("misc_f", "<f4"), ("grid_dims", "<u4"),
    ("off", "<u4"), ("off2", "<u4"), ("frame", "<u4"), ("misc_u", "<u4"),
    ("see_through", "<u4"), ("emitter", "<u4"), ("sun_su", "<f4"), ("sun_sv", "<f4"),
    ("sun_sprite", "<f4"), ("sky_map", "<f4"), ("water_plane", "<f4"),
]
# Params (scene.rs:157-228), 432 bytes = 27 vec4.
PARAMS = np.dtype([(name, t, 4) for name, t in _PARAM_FIELDS])
# lw_hit: trace's Hit (trace.rs:11-22) + final cell + hit flag.
HIT = np.dtype([("t", "<f4"), ("quad", "<u4"), ("texel", "<u4"), ("s", "<f4"), ("tt", "<f4"),
                ("cell", "<i4", 3), ("hit", "<u4"), ("pad", "<u4")])

# lw_view: one view of a batch.
VIEW = np.dtype([("params", PARAMS), ("spp", "<u4"), ("pad", "<u4", 3)])

assert QUAD.itemsize == 80 and MODEL.itemsize == 16 and EMITTER.itemsize == 32
assert VIEW.itemsize == 448
assert PARAMS.itemsize == 432 and HIT.itemsize == 40


def f32_bits(u: int) -> np.float32:
    """A u32 bit pattern as the f32 that holds it (Rust ``f32::from_bits``)."""
    return np.array([u], dtype="<u4").view("<f4")[0]


def bits_f32(f) -> int:
    """The bit pattern of an f32 (Rust ``f32::to_bits``)."""
    return int(np.array([f], dtype="<f4").view("<u4")[0])


def new_params() -> np.ndarray:
    """A zeroed Params record (``Params::default()``)."""
    return np.zeros((), dtype=PARAMS)


class SceneViewC(ctypes.Structure):
    """lw_scene_view: SceneView plus slice lengths (scene.rs:231-259)."""

    _fields_ = [
        ("params", ctypes.c_void_p),
        ("grid", ctypes.c_void_p), ("grid_len", ctypes.c_size_t),
        ("models", ctypes.c_void_p), ("models_len", ctypes.c_size_t),
        ("quads", ctypes.c_void_p), ("quads_len", ctypes.c_size_t),
        ("atlas", ctypes.c_void_p), ("atlas_len", ctypes.c_size_t),
        ("emitters", ctypes.c_void_p), ("emitters_len", ctypes.c_size_t),
        ("emitter_grid", ctypes.c_void_p), ("emitter_grid_len", ctypes.c_size_t),
        ("sky", ctypes.c_void_p), ("sky_len", ctypes.c_size_t),
    ]


NO_GROUP = 0xFFFFFFFF


class DiffDescC(ctypes.Structure):
    """lw_diff_desc: the parameters of a differentiable render."""

    _fields_ = [
        ("n_palette", ctypes.c_uint32), ("n_intensity", ctypes.c_uint32),
        ("quad_palette", ctypes.c_void_p), ("quad_intensity", ctypes.c_void_p),
        ("palette", ctypes.c_void_p), ("intensity", ctypes.c_void_p),
        ("sky_params", ctypes.c_uint32), ("rr_off", ctypes.c_uint32),
        ("d_sky_cache", ctypes.c_void_p),
        ("d_sun_dir", (ctypes.c_float * 3) * 3),
        ("d_sun_color", (ctypes.c_float * 3) * 3),
        ("d_sky_light", ctypes.c_float * 3),
        ("pad", ctypes.c_uint32),
    ]


class FitOptionsC(ctypes.Structure):
    """lw_fit_options: Adam's step size and moment decays, its epsilon, the factors' floor."""

    _fields_ = [("lr", ctypes.c_float), ("beta1", ctypes.c_float), ("beta2", ctypes.c_float),
                ("eps", ctypes.c_float), ("min_factor", ctypes.c_float),
                ("pad", ctypes.c_uint32 * 3)]


class FitStatsC(ctypes.Structure):
    """lw_fit_stats."""

    _fields_ = [("loss", ctypes.c_double), ("forward_ms", ctypes.c_double),
                ("adjoint_ms", ctypes.c_double), ("wall_ms", ctypes.c_double),
                ("samples", ctypes.c_uint64), ("step", ctypes.c_uint32), ("pad", ctypes.c_uint32)]


class TuneEntryC(ctypes.Structure):
    _fields_ = [("form", ctypes.c_uint32), ("waves", ctypes.c_uint32),
                ("per_lane", ctypes.c_uint32), ("pad", ctypes.c_uint32), ("ms", ctypes.c_double)]


class TuneResultC(ctypes.Structure):
    """lw_tune_result."""

    _fields_ = [("entries", TuneEntryC * 16), ("n", ctypes.c_uint32), ("best", ctypes.c_uint32),
                ("best_tiles", ctypes.c_uint32), ("pad", ctypes.c_uint32)]


FORMS = {0: "persistent", 1: "direct", 2: "wavefront", 3: "tiles"}


class TimingC(ctypes.Structure):
    """lw_timing."""

    _fields_ = [("trace_ms", ctypes.c_double), ("accumulate_ms", ctypes.c_double),
                ("wall_ms", ctypes.c_double), ("samples", ctypes.c_uint64),
                ("launches", ctypes.c_uint32), ("pad", ctypes.c_uint32)]


def _ptr(a: Optional[np.ndarray]) -> Optional[int]:
    if a is None or a.size == 0:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data


MODEL_OCCLUDES, MODEL_EMITS = 1, 2  # lw_voxel_scene.model_flags
SCENE_GRID, SCENE_EMITTERS, SCENE_EMITTER_GRID, SCENE_QUADS, SCENE_OCTANT_BOXES, SCENE_CELL_CUBES = (
    0, 1, 2, 3, 4, 5)  # lw_hip_scene_read


class VoxelSceneC(ctypes.Structure):
    _fields_ = [("dims", ctypes.c_uint32 * 3), ("see_through_count", ctypes.c_uint32),
                ("voxels", ctypes.c_void_p), ("brick_occludes", ctypes.c_void_p),
                ("models", ctypes.c_void_p), ("models_len", ctypes.c_size_t),
                ("model_flags", ctypes.c_void_p), ("quads", ctypes.c_void_p),
                ("quads_len", ctypes.c_size_t), ("entity_quads", ctypes.c_void_p),
                ("entity_quads_len", ctypes.c_size_t), ("atlas", ctypes.c_void_p),
                ("atlas_len", ctypes.c_size_t), ("atlas_size", ctypes.c_uint32 * 2),
                ("has_sun_sprite", ctypes.c_uint32), ("sun_sprite", ctypes.c_uint32 * 4),
                ("sky", ctypes.c_void_p), ("sky_len", ctypes.c_size_t)]


class GameSkyC(ctypes.Structure):
    """lw_game_sky: GameSkyMaps (sky.rs:885-905) without its buffer."""
    _fields_ = [("sizes", ctypes.c_float * 4), ("light", ctypes.c_float * 3),
                ("mean", ctypes.c_float * 3), ("scale", ctypes.c_float), ("veil", ctypes.c_float * 3),
                ("kernel_ms", ctypes.c_double), ("setup_ms", ctypes.c_double)]


class BuildStatsC(ctypes.Structure):
    _fields_ = [(k, ctypes.c_uint64) for k in (
        "bricks", "uniform_bricks", "buried_bricks", "instances", "overlay_bricks", "entity_quads",
        "emitters", "grid_len", "quads_len", "emitters_len", "emitter_grid_len")] + [
        ("kernel_ms", ctypes.c_double), ("wall_ms", ctypes.c_double)]


@dataclass
class VoxelVolume:
    """What build_hq (crates/lumenwright/src/scene/voxel.rs:111-491) holds once it has
    interned the view's block-model instances: one model id per voxel (4096 per brick, bricks
    in y, z, x order, voxels (ly * 16 + lz) * 16 + lx), the models with their quads and
    LW_MODEL_* flags, the entity overlay's quads and the atlas (lw_voxel_scene)."""

    dims: tuple
    voxels: np.ndarray
    models: np.ndarray
    model_flags: np.ndarray
    quads: np.ndarray
    atlas: np.ndarray
    atlas_size: tuple
    brick_occludes: Optional[np.ndarray] = None
    entity_quads: np.ndarray = field(default_factory=lambda: np.zeros(0, dtype=QUAD))
    see_through_count: int = 0
    sun_sprite: Optional[tuple] = None

    def c(self, sky=None):
        """The lw_voxel_scene and the arrays it points into (keep both alive)."""
        keep = {
            "voxels": np.ascontiguousarray(self.voxels, dtype="<u2").reshape(-1),
            "models": np.ascontiguousarray(self.models, dtype=MODEL),
            "flags": np.ascontiguousarray(self.model_flags, dtype=np.uint8),
            "quads": np.ascontiguousarray(self.quads, dtype=QUAD),
            "entity": np.ascontiguousarray(self.entity_quads, dtype=QUAD),
            "atlas": np.ascontiguousarray(self.atlas, dtype="<u4"),
            "sky": np.ascontiguousarray(sky if sky is not None else np.zeros(0), dtype="<f4"),
        }
        occ = None
        if self.brick_occludes is not None:
            occ = keep["occ"] = np.ascontiguousarray(self.brick_occludes, dtype=np.uint8)
        c = VoxelSceneC()
        c.dims[:] = [int(d) for d in self.dims]
        c.see_through_count = self.see_through_count
        c.voxels = _ptr(keep["voxels"])
        c.brick_occludes = _ptr(occ) if occ is not None else None
        c.models, c.models_len = _ptr(keep["models"]), keep["models"].size
        c.model_flags = _ptr(keep["flags"])
        c.quads, c.quads_len = _ptr(keep["quads"]), keep["quads"].size
        c.entity_quads, c.entity_quads_len = _ptr(keep["entity"]), keep["entity"].size
        c.atlas, c.atlas_len = _ptr(keep["atlas"]), keep["atlas"].size
        c.atlas_size[:] = [int(a) for a in self.atlas_size]
        if self.sun_sprite is not None:
            c.has_sun_sprite = 1
            c.sun_sprite[:] = [int(a) for a in self.sun_sprite]
        c.sky, c.sky_len = _ptr(keep["sky"]), keep["sky"].size
        return c, keep


@dataclass
class HqScene:
    """The HQ scene buffers (crates/lumenwright/src/scene/voxel.rs:17-49)."""

    grid: np.ndarray
    models: np.ndarray
    quads: np.ndarray
    atlas: np.ndarray
    atlas_size: tuple
    dims: tuple
    offsets: tuple
    see_through: tuple = (0, 0)
    emitters: np.ndarray = field(default_factory=lambda: np.zeros(1, dtype=EMITTER))
    emitter_count: int = 0
    emitter_grid: np.ndarray = field(default_factory=lambda: np.zeros(8, dtype="<u4"))
    sun_sprite: Optional[tuple] = None

    def memory_bytes(self) -> int:
        """voxel.rs:91-99"""
        return (self.grid.nbytes + self.models.nbytes + self.quads.nbytes + self.atlas.nbytes
                + self.emitters.nbytes + self.emitter_grid.nbytes)


@dataclass
class HqFrame:
    """HqFrame (crates/lumenwright/src/render/mod.rs:80-89)."""

    scene: HqScene
    params: np.ndarray
    spp: int
    sky: np.ndarray

    def view(self) -> "BoundView":
        return BoundView(self.params, self.scene.grid, self.scene.models, self.scene.quads,
                         self.scene.atlas, self.scene.emitters, self.scene.emitter_grid,
                         self.sky)


class BoundView:
    """A ctypes lw_scene_view over numpy buffers, keeping the buffers alive."""

    def __init__(self, params, grid, models, quads, atlas, emitters, emitter_grid, sky):
        self.params = np.ascontiguousarray(params, dtype=PARAMS).reshape(())
        self.grid = np.ascontiguousarray(grid, dtype="<u4")
        self.models = np.ascontiguousarray(models, dtype=MODEL)
        self.quads = np.ascontiguousarray(quads, dtype=QUAD)
        self.atlas = np.ascontiguousarray(atlas, dtype="<u4")
        self.emitters = np.ascontiguousarray(emitters, dtype=EMITTER)
        self.emitter_grid = np.ascontiguousarray(emitter_grid, dtype="<u4")
        self.sky = np.ascontiguousarray(sky if sky is not None else np.zeros(0), dtype="<f4")
        self.c = SceneViewC(
            self.params.ctypes.data,
            _ptr(self.grid), self.grid.size, _ptr(self.models), self.models.size,
            _ptr(self.quads), self.quads.size, _ptr(self.atlas), self.atlas.size,
            _ptr(self.emitters), self.emitters.size, _ptr(self.emitter_grid),
            self.emitter_grid.size, _ptr(self.sky), self.sky.size)

    @property
    def ref(self):
        return ctypes.byref(self.c)


# ---- quick mode (lw_quick_vertex, lw_quick_frame) ----

# Vertex (crates/lumenwright/src/scene/mesh.rs:12-21): camera-relative position, atlas UV,
# sRGB colour.
QUICK_VERTEX = np.dtype([("pos", "<f4", 3), ("uv", "<f4", 2), ("color", "u1", 4)])
assert QUICK_VERTEX.itemsize == 24


class QuickFrameC(ctypes.Structure):
    """lw_quick_frame."""
    _fields_ = [("vertices", ctypes.c_void_p), ("vertices_len", ctypes.c_size_t),
                ("opaque", ctypes.c_void_p), ("opaque_len", ctypes.c_size_t),
                ("translucent", ctypes.c_void_p), ("translucent_len", ctypes.c_size_t),
                ("see_through", ctypes.c_void_p), ("see_through_len", ctypes.c_size_t),
                ("atlas_pixels", ctypes.c_void_p), ("atlas_size", ctypes.c_uint32 * 2),
                ("atlas_mips", ctypes.c_void_p), ("atlas_mips_len", ctypes.c_size_t),
                ("mip_levels", ctypes.c_uint32), ("mip_rows", ctypes.c_uint32),
                ("stars", ctypes.c_void_p), ("stars_len", ctypes.c_size_t),
                ("width", ctypes.c_uint32), ("height", ctypes.c_uint32),
                ("view_proj", ctypes.c_float * 16), ("sky", ctypes.c_void_p)]


@dataclass
class QuickMesh:
    """Mesh (crates/lumenwright/src/scene/mesh.rs:24-36): the vertices and three index lists,
    three indices a triangle (opaque and cutout, translucent back to front, see-through)."""

    vertices: np.ndarray
    opaque: np.ndarray
    translucent: np.ndarray = field(default_factory=lambda: np.zeros(0, "<u4"))
    see_through: np.ndarray = field(default_factory=lambda: np.zeros(0, "<u4"))

    def triangle_count(self) -> int:
        return (self.opaque.size + self.translucent.size + self.see_through.size) // 3


@dataclass
class QuickFrame:
    """QuickFrame (crates/lumenwright/src/render/mod.rs:53-78). view_proj is glam's Mat4 as 16
    floats, column-major; sky is a QuickSky (lumenwright_amd.frames.QuickSkyC)."""

    mesh: QuickMesh
    atlas_pixels: np.ndarray
    atlas_size: tuple
    atlas_mips: np.ndarray
    mip_levels: int
    mip_rows: int
    stars: np.ndarray
    width: int
    height: int
    view_proj: np.ndarray
    sky: object

    def c(self):
        """The lw_quick_frame and the arrays it points into (keep both alive)."""
        keep = {
            "v": np.ascontiguousarray(self.mesh.vertices, dtype=QUICK_VERTEX),
            "o": np.ascontiguousarray(self.mesh.opaque, dtype="<u4"),
            "t": np.ascontiguousarray(self.mesh.translucent, dtype="<u4"),
            "s": np.ascontiguousarray(self.mesh.see_through, dtype="<u4"),
            "atlas": np.ascontiguousarray(self.atlas_pixels, dtype="<u4"),
            "mips": np.ascontiguousarray(self.atlas_mips, dtype="<u4"),
            "stars": np.ascontiguousarray(self.stars, dtype="<u4"),
            "sky": self.sky,
        }
        c = QuickFrameC()
        c.vertices, c.vertices_len = _ptr(keep["v"]), keep["v"].size
        c.opaque, c.opaque_len = _ptr(keep["o"]), keep["o"].size
        c.translucent, c.translucent_len = _ptr(keep["t"]), keep["t"].size
        c.see_through, c.see_through_len = _ptr(keep["s"]), keep["s"].size
        c.atlas_pixels = _ptr(keep["atlas"])
        c.atlas_size[:] = [int(a) for a in self.atlas_size]
        c.atlas_mips, c.atlas_mips_len = _ptr(keep["mips"]), keep["mips"].size
        c.mip_levels, c.mip_rows = int(self.mip_levels), int(self.mip_rows)
        c.stars, c.stars_len = _ptr(keep["stars"]), keep["stars"].size
        c.width, c.height = int(self.width), int(self.height)
        c.view_proj[:] = [float(x) for x in np.asarray(self.view_proj, "<f4").reshape(16)]
        c.sky = ctypes.addressof(self.sky)
        return c, keep
