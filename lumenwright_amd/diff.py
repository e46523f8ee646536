"""Differentiable rendering on the device: parameters, sky derivative tables, the loss.

The reference renders with fixed materials and a fixed sky; this module describes the
parameters lw_hip_render_diff and lw_hip_adjoint (include/lw/abi.h) take:

  * palette groups: an RGB factor on the albedo of every quad in the group (a block type);
  * intensity groups: a factor on the emittance of every quad in the group (an emitter);
  * the sky: Preetham turbidity (the reference fixes 2.5, sky.rs:667) and the sun's
    altitude and azimuth in radians (sky.rs:41-44).

The sky enters the device through host-made tables: the sky cache and the Params' sun
fields are functions of the three sky parameters through the host sky setup (Preetham::new,
hq_sky, build_sky_cache); their derivatives are taken here by central differences of that
setup, with the same step the finite-difference checks use.
"""

from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass, field, replace
from typing import List, Optional

import numpy as np

from . import frames
from .abi import NO_GROUP, DiffDescC

SKY_NAMES = ("turbidity", "sun_altitude", "sun_azimuth")


@dataclass
class SkyDerivs:
    """d(sky cache, sun_dir, sun_color, sky light) / d(turbidity, altitude, azimuth)."""

    cache: np.ndarray       # (3, cache_len) f32
    sun_dir: np.ndarray     # (3, 3)
    sun_color: np.ndarray   # (3, 3)
    sky_light: np.ndarray   # (3,)


@dataclass
class DiffParams:
    """The parameters of a differentiable render of one world."""

    quad_palette: np.ndarray             # (quads,) u32: group or NO_GROUP
    quad_intensity: np.ndarray           # (quads,) u32
    palette: np.ndarray                  # (n_palette, 3) f32, > 0
    intensity: np.ndarray                # (n_intensity,) f32, > 0
    sky: Optional[np.ndarray] = None     # (turbidity, altitude, azimuth); None: the config's
    rr_off: bool = False
    _keep: List = field(default_factory=list, repr=False)

    @property
    def n_palette(self) -> int:
        return int(self.palette.shape[0])

    @property
    def n_intensity(self) -> int:
        return int(self.intensity.shape[0])

    @property
    def n_grads(self) -> int:
        return 3 * self.n_palette + self.n_intensity + 3

    def with_values(self, palette=None, intensity=None, sky=None) -> "DiffParams":
        return replace(self,
                       palette=self.palette if palette is None else np.asarray(palette, "<f4").reshape(-1, 3),
                       intensity=self.intensity if intensity is None else np.asarray(intensity, "<f4"),
                       sky=self.sky if sky is None else np.asarray(sky, "<f8"), _keep=[])

    def vector(self) -> np.ndarray:
        """palette RGB, intensities, then the sky triple, as one f64 vector."""
        sky = self.sky if self.sky is not None else np.zeros(3)
        return np.concatenate([self.palette.astype("<f8").ravel(), self.intensity.astype("<f8"),
                               np.asarray(sky, "<f8")])

    def desc(self, derivs: Optional[SkyDerivs] = None) -> DiffDescC:
        """The lw_diff_desc of these parameters; the arrays it points to live with self."""
        arrays = [np.ascontiguousarray(self.quad_palette, "<u4"),
                  np.ascontiguousarray(self.quad_intensity, "<u4"),
                  np.ascontiguousarray(self.palette, "<f4").ravel(),
                  np.ascontiguousarray(self.intensity, "<f4")]
        d = DiffDescC()
        d.n_palette, d.n_intensity = self.n_palette, self.n_intensity
        d.quad_palette, d.quad_intensity, d.palette, d.intensity = [a.ctypes.data for a in arrays]
        d.rr_off = 1 if self.rr_off else 0
        if derivs is not None:
            cache = np.ascontiguousarray(derivs.cache, "<f4")
            arrays.append(cache)
            d.sky_params = 1
            d.d_sky_cache = cache.ctypes.data
            for k in range(3):
                for c in range(3):
                    d.d_sun_dir[k][c] = float(derivs.sun_dir[k, c])
                    d.d_sun_color[k][c] = float(derivs.sun_color[k, c])
                d.d_sky_light[k] = float(derivs.sky_light[k])
        self._keep[:] = arrays
        return d


def sky_config(cfg: frames.FrameConfig, sky) -> frames.FrameConfig:
    """cfg with the sky triple (turbidity, altitude, azimuth) in place."""
    t, alt, az = (float(x) for x in sky)
    return replace(cfg, turbidity=t, sun_altitude=alt, sun_azimuth=az)


def base_sky(cfg: frames.FrameConfig) -> np.ndarray:
    if cfg.sun_altitude is None or cfg.sun_azimuth is None:
        raise ValueError("the sky parameters need a sun given by altitude and azimuth")
    return np.array([cfg.turbidity, cfg.sun_altitude, cfg.sun_azimuth], dtype="<f8")


def frame_at(cfg: frames.FrameConfig, scene, sky) -> frames.HqFrame:
    """The frame (Params and sky cache) of cfg's camera with the sky triple `sky`."""
    return frames.make_frame(sky_config(cfg, sky), scene)


def sky_derivatives(cfg: frames.FrameConfig, scene, sky, h: float = 1e-3) -> SkyDerivs:
    """Central differences of the host sky setup in each sky parameter, step h."""
    sky = np.asarray(sky, dtype="<f8")
    caches, sdir, scol, slight = [], [], [], []
    for k in range(3):
        e = np.zeros(3)
        e[k] = h
        fp = frame_at(cfg, scene, sky + e)
        fm = frame_at(cfg, scene, sky - e)
        dc = fp.sky.astype("<f8") - fm.sky.astype("<f8")
        # Hue is an angle in turns: rgb_to_hsl (integrator.rs:79-104) gives -1/6 on one side
        # of r = b and 5/6 on the other, the same hue. A node that crosses there within the
        # step has the circular difference as its hue's change.
        hue = dc[0::3]
        dc[0::3] = hue - np.round(hue)
        caches.append(dc / (2 * h))
        pp, pm = fp.params, fm.params
        sdir.append((pp["sun_dir"][:3].astype("<f8") - pm["sun_dir"][:3]) / (2 * h))
        scol.append((pp["sun_color"][:3].astype("<f8") - pm["sun_color"][:3]) / (2 * h))
        slight.append((float(pp["sky_sun"][3]) - float(pm["sky_sun"][3])) / (2 * h))
    return SkyDerivs(np.stack(caches).astype("<f4"), np.stack(sdir), np.stack(scol),
                     np.asarray(slight))


def block_type_groups(scene, n_types: Optional[int] = None) -> np.ndarray:
    """Palette groups by block type (the world builder's per-quad type)."""
    g = np.asarray(scene.quad_type, "<u4").copy()
    if n_types is not None:
        g[(g != NO_GROUP) & (g >= n_types)] = NO_GROUP
    return g


def emitter_groups(scene, n_groups: Optional[int] = None, seed: int = 0) -> np.ndarray:
    """Intensity groups: one per emitter, or each emitter in one of n_groups seeded groups."""
    e = np.asarray(scene.quad_emitter, "<u4").copy()
    if n_groups is None:
        return e
    n_em = int(e[e != NO_GROUP].max()) + 1 if np.any(e != NO_GROUP) else 0
    assign = np.random.default_rng(seed).integers(0, n_groups, size=max(n_em, 1)).astype("<u4")
    out = np.full_like(e, NO_GROUP)
    m = e != NO_GROUP
    out[m] = assign[e[m]]
    return out


def identity(quad_palette, quad_intensity, n_palette: int, n_intensity: int, sky=None,
             rr_off: bool = False) -> DiffParams:
    return DiffParams(np.asarray(quad_palette, "<u4"), np.asarray(quad_intensity, "<u4"),
                      np.ones((n_palette, 3), "<f4"), np.ones(n_intensity, "<f4"),
                      None if sky is None else np.asarray(sky, "<f8"), rr_off)


def l2_loss(image: np.ndarray, target: np.ndarray):
    """L = mean((I - T)^2) over pixels and channels, and dL/dI (f32, w*h*3)."""
    diff = image.astype("<f8") - target.astype("<f8")
    n = diff.size
    return float(np.mean(diff * diff)), np.ascontiguousarray((2.0 / n) * diff, "<f4")


@dataclass
class Problem:
    """A differentiable-rendering problem: a world, a camera, its parameters at the start
    and at the target."""

    cfg: frames.FrameConfig
    scene: object
    start: DiffParams
    target: DiffParams


def config3_problem(cfg: Optional[frames.FrameConfig] = None, seed: int = 7) -> Problem:
    """cfg3: 32 block-type palettes (96 albedo factors), 16 emitter intensities, turbidity
    and the sun's two angles. The target perturbs every albedo factor by U(-0.1, 0.1), every
    intensity by U(-20%, 20%) and sets turbidity to 3.5."""
    cfg = cfg or frames.CONFIGS[3]
    scene = frames.build_world(cfg.world)
    qp = block_type_groups(scene, cfg.world.cube_types)
    qi = emitter_groups(scene)
    n_int = int(qi[qi != NO_GROUP].max()) + 1 if np.any(qi != NO_GROUP) else 0
    start = identity(qp, qi, cfg.world.cube_types, n_int, base_sky(cfg))
    rng = np.random.default_rng(seed)
    pal = (1.0 + rng.uniform(-0.1, 0.1, size=(cfg.world.cube_types, 3))).astype("<f4")
    inten = (1.0 + rng.uniform(-0.2, 0.2, size=n_int)).astype("<f4")
    sky = base_sky(cfg).copy()
    sky[0] = 3.5
    return Problem(cfg, scene, start, start.with_values(pal, inten, sky))


@dataclass
class BatchProblem:
    """An inverse fit over a batch of views: the views, their parameter groups, the start and
    the seeded ground truth the targets are rendered with."""

    cfg: frames.FrameConfig
    scene: object
    base: frames.HqFrame
    views: list
    start: DiffParams
    truth: DiffParams


def config4_problem(n_views: int = 32, size=(640, 360), spp: int = 128, max_bounces: int = 6,
                    n_groups: int = 128, seed: int = 4, world=None) -> BatchProblem:
    """cfg4: seeded views of the cfg1 world; 64 block-type palettes (192 albedo factors) and
    the emitters in 128 seeded intensity groups. Ground truth: albedo factors U(0.8, 1.2),
    intensities U(0.6, 1.4); the fit starts from 1."""
    cfg = frames.CONFIGS[1] if world is None else replace(frames.CONFIGS[1], world=world)
    scene = frames.build_world(cfg.world)
    base, views = frames.batch_views(cfg, scene, n_views=n_views, seed=seed, sizes=(size,),
                                     spps=(spp,), max_bounces=max_bounces)
    n_types = cfg.world.cube_types + cfg.world.cutout_types
    qp = block_type_groups(scene, n_types)
    qi = emitter_groups(scene, n_groups, seed=seed)
    start = identity(qp, qi, n_types, n_groups)
    rng = np.random.default_rng(seed + 100)
    truth = start.with_values(rng.uniform(0.8, 1.2, size=(n_types, 3)).astype("<f4"),
                              rng.uniform(0.6, 1.4, size=n_groups).astype("<f4"))
    return BatchProblem(cfg, scene, base, views, start, truth)


def parameter_error(p: DiffParams, truth: DiffParams, used_palette=None, used_intensity=None):
    """Mean absolute error of the palette and intensity factors against the truth, over the
    groups the views can see (all when the masks are None)."""
    dp = np.abs(p.palette - truth.palette)
    di = np.abs(p.intensity - truth.intensity)
    if used_palette is not None:
        dp = dp[used_palette]
    if used_intensity is not None:
        di = di[used_intensity]
    return float(dp.mean()) if dp.size else 0.0, float(di.mean()) if di.size else 0.0
