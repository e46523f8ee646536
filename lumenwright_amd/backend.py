"""The HIP backend behind the reference's Backend contract.

Mirrors ``trait Backend { name, render_quick, render_hq }``
(crates/lumenwright/src/render/mod.rs:92-103) and CudaBackend's construction and errors
(crates/lumenwright/src/render/cuda.rs:29-75). Every call goes through the C ABI of
include/lw/abi.h; there is no CPU path here.
"""

from __future__ import annotations

import ctypes
from typing import Callable, Optional

import numpy as np

from . import _native
from .abi import (EMITTER, FORMS, HIT, MODEL, PARAMS, QUAD, SCENE_CELL_CUBES, SCENE_EMITTER_GRID,
                  SCENE_EMITTERS, SCENE_GRID, SCENE_OCTANT_BOXES, SCENE_QUADS, VIEW, BoundView, BuildStatsC, FitOptionsC, FitStatsC,
                  GameSkyC, HqFrame, HqScene, QuickFrame, TimingC, TuneResultC, VoxelVolume)
from .frames import SceneInfoC

LW_OK, LW_EINVAL, LW_ENODEV, LW_EHIP, LW_ENOMEM, LW_EUNSUPPORTED = 0, -1, -2, -3, -4, -5


class BackendError(RuntimeError):
    """A failed backend call (the reference's anyhow::Error); `code` is the LW_E* value."""

    def __init__(self, code: int, message: str):
        super().__init__(message)
        self.code = code


def probe() -> str:
    """Describes device 0 or says why HIP is unavailable (render/cuda.rs:44-51)."""
    buf = ctypes.create_string_buffer(512)
    rc = _native.lib().lw_hip_probe(buf, 512)
    if rc != LW_OK:
        raise BackendError(rc, buf.value.decode())
    return buf.value.decode()


def _addr(a: np.ndarray):
    return ctypes.c_void_p(a.ctypes.data)


class HipScene:
    """A scene validated and resident in device memory (lw_hip_scene)."""

    def __init__(self, backend: "HipBackend", handle: int, view: Optional[BoundView] = None):
        self.backend = backend
        self.handle = handle
        self.params = view.params.copy() if view is not None else None

    def free(self):
        if self.handle:
            _native.lib().lw_hip_scene_free(self.backend.handle, self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class HipBackend:
    """BackendKind::Hip: the HQ path tracer on an AMD GPU."""

    def __init__(self, handle: int):
        self.handle = handle

    @classmethod
    def open(cls, device: int = 0) -> "HipBackend":
        """CudaBackend::new's counterpart (render/cuda.rs:55-62)."""
        L = _native.lib()
        h = ctypes.c_void_p()
        rc = L.lw_hip_open(device, ctypes.byref(h))
        if rc != LW_OK:
            raise BackendError(rc, "HIP: " + L.lw_last_error().decode())
        return cls(h.value)

    def _check(self, rc: int):
        if rc != LW_OK:
            raise BackendError(rc, _native.lib().lw_hip_last_error(self.handle).decode())

    def name(self) -> str:
        """Backend::name (render/mod.rs:94)."""
        return _native.lib().lw_hip_name(self.handle).decode()

    def render_quick(self, frame: QuickFrame) -> np.ndarray:
        """Backend::render_quick (render/mod.rs:96): the frame rasterized on the device as the
        reference's CPU rasterizer draws it (render/cpu/raster.rs): width * height packed RGBA8
        words (red in the low byte), row major, top row first. last_timing() then holds the
        device time (trace_ms, inputs resident) of the call."""
        c, keep = frame.c()
        out = np.zeros(int(frame.width) * int(frame.height), dtype="<u4")
        self._check(_native.lib().lw_hip_render_quick(self.handle, ctypes.byref(c), _addr(out)))
        del keep
        return out

    def render_hq(self, frame: HqFrame,
                  progress: Optional[Callable[[int, int], None]] = None) -> np.ndarray:
        """Backend::render_hq (render/mod.rs:98-102): the mean linear radiance per pixel,
        row-major, top row first, as (w*h, 3) float32."""
        view = frame.view()
        w, h = int(view.params["frame"][0]), int(view.params["frame"][1])
        out = np.zeros((w * h, 3), dtype="<f4")
        cb = _native.PROGRESS_FN((lambda d, t, u: progress(d, t)) if progress else 0)
        self._check(_native.lib().lw_hip_render_hq(self.handle, view.ref, frame.spp, cb, None,
                                                   _addr(out)))
        return out

    def upload(self, frame: HqFrame) -> HipScene:
        """Validates and uploads the frame's scene buffers once (lw_hip_scene_upload)."""
        view = frame.view()
        h = ctypes.c_void_p()
        self._check(_native.lib().lw_hip_scene_upload(self.handle, view.ref, ctypes.byref(h)))
        return HipScene(self, h.value, view)

    def build_scene(self, vol: VoxelVolume, sky: Optional[np.ndarray] = None):
        """lw_hip_scene_build: build_hq's layout (scene/voxel.rs:305-491) made on the device
        from an interned voxel volume. Returns (HipScene, HqScene-shaped info, stats): the
        info is an HqScene whose buffers are empty and whose layout facts (dims, offsets,
        see-through range, emitter count) are the built scene's, for hq_params."""
        c, keep = vol.c(sky)
        info, stats, h = SceneInfoC(), BuildStatsC(), ctypes.c_void_p()
        self._check(_native.lib().lw_hip_scene_build(self.handle, ctypes.byref(c), ctypes.byref(info),
                                                     ctypes.byref(stats), ctypes.byref(h)))
        del keep
        layout = HqScene(grid=np.zeros(0, "<u4"), models=np.zeros(int(info.models_len), MODEL),
                         quads=np.zeros(int(info.quads_len), QUAD), atlas=np.zeros(0, "<u4"),
                         atlas_size=tuple(info.atlas_size), dims=tuple(info.dims),
                         offsets=tuple(info.offsets), see_through=tuple(info.see_through),
                         emitter_count=int(info.emitter_count),
                         sun_sprite=tuple(info.sun_sprite) if info.has_sun_sprite else None)
        st = {k: getattr(stats, k) for k, _ in BuildStatsC._fields_}
        return HipScene(self, h.value), layout, st

    def scene_read(self, scene: HipScene, which: int) -> np.ndarray:
        """lw_hip_scene_read: one of a resident scene's buffers (SCENE_*) on the host."""
        L = _native.lib()
        n = ctypes.c_size_t()
        self._check(L.lw_hip_scene_read(self.handle, scene.handle, which, None, 0, ctypes.byref(n)))
        dtype = {SCENE_GRID: "<u4", SCENE_EMITTER_GRID: "<u4", SCENE_EMITTERS: EMITTER,
                 SCENE_QUADS: QUAD, SCENE_OCTANT_BOXES: "u1", SCENE_CELL_CUBES: "<u4"}[which]
        out = np.zeros(n.value, dtype=dtype)
        self._check(L.lw_hip_scene_read(self.handle, scene.handle, which, _addr(out), out.size,
                                        ctypes.byref(n)))
        return out

    def render_hq_scene(self, scene: HipScene, params: np.ndarray, spp: int,
                        progress: Optional[Callable[[int, int], None]] = None) -> np.ndarray:
        """render_hq over a resident scene."""
        p = np.ascontiguousarray(params, dtype=PARAMS).reshape(())
        w, h = int(p["frame"][0]), int(p["frame"][1])
        out = np.zeros((w * h, 3), dtype="<f4")
        cb = _native.PROGRESS_FN((lambda d, t, u: progress(d, t)) if progress else 0)
        self._check(_native.lib().lw_hip_render_hq_scene(self.handle, scene.handle, _addr(p), spp,
                                                         cb, None, _addr(out)))
        return out

    def render_hq_scene_noalloc(self, scene: HipScene, params: np.ndarray, spp: int,
                                progress: Optional[Callable[[int, int], None]] = None) -> np.ndarray:
        """lw_hip_render_hq_scene_noalloc: render_hq over a resident scene in the library's
        fixed scratch memory, allocating nothing."""
        p = np.ascontiguousarray(params, dtype=PARAMS).reshape(())
        w, h = int(p["frame"][0]), int(p["frame"][1])
        out = np.zeros((w * h, 3), dtype="<f4")
        cb = _native.PROGRESS_FN((lambda d, t, u: progress(d, t)) if progress else 0)
        self._check(_native.lib().lw_hip_render_hq_scene_noalloc(
            self.handle, scene.handle, _addr(p), spp, cb, None, _addr(out)))
        return out

    def count_traces(self, scene: HipScene, params: np.ndarray, spp: int) -> int:
        """lw_hip_count_traces: the trace calls a render_hq_scene of (params, spp) makes,
        counted on the device by an untimed replay of its paths."""
        p = np.ascontiguousarray(params, dtype=PARAMS).reshape(())
        n = ctypes.c_uint64()
        self._check(_native.lib().lw_hip_count_traces(self.handle, scene.handle, _addr(p), spp,
                                                      ctypes.byref(n)))
        return int(n.value)

    def render_hq_batch(self, scene: HipScene, views,
                        progress: Optional[Callable[[int, int], None]] = None) -> list:
        """Renders [(params, spp), ...] of one resident scene as one batch
        (lw_hip_render_hq_batch); one (w*h, 3) float32 image per view."""
        arr = np.zeros(len(views), dtype=VIEW)
        outs = []
        for i, (params, spp) in enumerate(views):
            arr[i]["params"] = params
            arr[i]["spp"] = spp
            w, h = int(params["frame"][0]), int(params["frame"][1])
            outs.append(np.zeros((w * h, 3), dtype="<f4"))
        ptrs = (ctypes.c_void_p * len(views))(*[o.ctypes.data for o in outs])
        cb = _native.PROGRESS_FN((lambda d, t, u: progress(d, t)) if progress else 0)
        self._check(_native.lib().lw_hip_render_hq_batch(self.handle, scene.handle, _addr(arr),
                                                         len(views), cb, None, ptrs))
        return outs

    def render_diff(self, scene: HipScene, params: np.ndarray, spp: int, desc,
                    f64: bool = False):
        """lw_hip_render_diff: the image under a lw_diff_desc's parameters, (w*h, 3) f32; with
        f64, also the means before their rounding to f32."""
        p = np.ascontiguousarray(params, dtype=PARAMS).reshape(())
        w, h = int(p["frame"][0]), int(p["frame"][1])
        out = np.zeros((w * h, 3), dtype="<f4")
        out64 = np.zeros((w * h, 3), dtype="<f8") if f64 else None
        self._check(_native.lib().lw_hip_render_diff(
            self.handle, scene.handle, _addr(p), spp, ctypes.byref(desc), _addr(out),
            _addr(out64) if f64 else None))
        return (out, out64) if f64 else out

    def _views(self, views):
        arr = np.zeros(len(views), dtype=VIEW)
        for i, (params, spp) in enumerate(views):
            arr[i]["params"] = params
            arr[i]["spp"] = spp
        return arr

    def render_diff_batch(self, scene: HipScene, views, desc) -> list:
        """lw_hip_render_diff_batch over [(params, spp), ...]; one (w*h, 3) f32 per view."""
        arr = self._views(views)
        outs = [np.zeros((int(p["frame"][0]) * int(p["frame"][1]), 3), dtype="<f4") for p, _ in views]
        ptrs = (ctypes.c_void_p * len(views))(*[o.ctypes.data for o in outs])
        self._check(_native.lib().lw_hip_render_diff_batch(self.handle, scene.handle, _addr(arr),
                                                           len(views), ctypes.byref(desc), ptrs))
        return outs

    @staticmethod
    def _per_view_images(views, images, what: str):
        """Each view's w*h*3 array as contiguous f32, checked against the view's size: the
        library copies w*h*3 floats from every pointer."""
        if len(images) != len(views):
            raise ValueError(f"{len(images)} {what} arrays for {len(views)} views")
        out = []
        for i, ((params, _), a) in enumerate(zip(views, images)):
            a = np.ascontiguousarray(a, dtype="<f4").reshape(-1)
            want = int(params["frame"][0]) * int(params["frame"][1]) * 3
            if a.size != want:
                raise ValueError(f"view {i}: {what} has {a.size} values, the view needs {want}")
            out.append(a)
        return out

    def adjoint_batch(self, scene: HipScene, views, desc, n_grads: int, dldis) -> np.ndarray:
        """lw_hip_adjoint_batch: the gradient of sum over views of dldis[i] . I_i."""
        arr = self._views(views)
        ds = self._per_view_images(views, dldis, "dL/dI")
        ptrs = (ctypes.c_void_p * len(views))(*[d.ctypes.data for d in ds])
        out = np.zeros(n_grads, dtype="<f8")
        self._check(_native.lib().lw_hip_adjoint_batch(self.handle, scene.handle, _addr(arr),
                                                       len(views), ctypes.byref(desc), ptrs,
                                                       _addr(out)))
        return out

    def fit(self, scene: HipScene, views, desc, targets, lr: float = 1e-2, beta1: float = 0.9,
            beta2: float = 0.999, eps: float = 1e-8, min_factor: float = 1e-3) -> "HipFit":
        """lw_hip_fit_new: an inverse fit of desc's factors to the views' target images."""
        arr = self._views(views)
        ts = self._per_view_images(views, targets, "target")
        ptrs = (ctypes.c_void_p * len(views))(*[t.ctypes.data for t in ts])
        opt = FitOptionsC(lr, beta1, beta2, eps, min_factor)
        h = ctypes.c_void_p()
        self._check(_native.lib().lw_hip_fit_new(self.handle, scene.handle, _addr(arr), len(views),
                                                 ctypes.byref(desc), ptrs, ctypes.byref(opt),
                                                 ctypes.byref(h)))
        return HipFit(self, scene, h.value, desc.n_palette, desc.n_intensity)

    def adjoint(self, scene: HipScene, params: np.ndarray, spp: int, desc, n_grads: int,
                d_loss_d_image: np.ndarray) -> np.ndarray:
        """lw_hip_adjoint: sum over pixels of dL/dI . dI/dtheta, (n_grads,) f64."""
        p = np.ascontiguousarray(params, dtype=PARAMS).reshape(())
        g = np.ascontiguousarray(d_loss_d_image, dtype="<f4").reshape(-1)
        if g.size != int(p["frame"][0]) * int(p["frame"][1]) * 3:
            raise ValueError(f"dL/dI has {g.size} values, the frame needs "
                             f"{int(p['frame'][0]) * int(p['frame'][1]) * 3}")
        out = np.zeros(n_grads, dtype="<f8")
        self._check(_native.lib().lw_hip_adjoint(self.handle, scene.handle, _addr(p), spp,
                                                 ctypes.byref(desc), _addr(g), _addr(out)))
        return out

    def autotune(self, scene: HipScene, params: np.ndarray, probe_spp: int = 4) -> dict:
        """lw_hip_autotune: times each launch shape on a probe of the frame and keeps the
        fastest on this handle; returns the timings and the choice."""
        p = np.ascontiguousarray(params, dtype=PARAMS).reshape(())
        r = TuneResultC()
        self._check(_native.lib().lw_hip_autotune(self.handle, scene.handle, _addr(p), probe_spp,
                                                  ctypes.byref(r)))
        rows = [{"form": FORMS[e.form], "waves": e.waves, "per_lane": e.per_lane, "ms": e.ms}
                for e in r.entries[:r.n]]
        out = {"candidates": rows, "best": rows[r.best]}
        if r.best_tiles != 0xFFFFFFFF:
            out["best_tiles"] = rows[r.best_tiles]
        return out

    def tonemap(self, hdr: np.ndarray, exposure: float = 1.0, op: int = 0) -> np.ndarray:
        """lw_hip_tonemap (render/post.rs:148-170): (n, 3) linear to n packed RGBA8."""
        h = np.ascontiguousarray(hdr, dtype="<f4").reshape(-1, 3)
        out = np.zeros(h.shape[0], dtype="<u4")
        self._check(_native.lib().lw_hip_tonemap(self.handle, _addr(h), h.shape[0], exposure, op,
                                                 _addr(out)))
        return out

    def tonemap_mix(self, hdr: np.ndarray, exposure: float, op_a: int, op_b: int,
                    weight: float) -> np.ndarray:
        """lw_hip_tonemap_mix (render/post.rs:176-215)."""
        h = np.ascontiguousarray(hdr, dtype="<f4").reshape(-1, 3)
        out = np.zeros(h.shape[0], dtype="<u4")
        self._check(_native.lib().lw_hip_tonemap_mix(self.handle, _addr(h), h.shape[0], exposure,
                                                     op_a, op_b, weight, _addr(out)))
        return out

    def encode(self, x: np.ndarray) -> np.ndarray:
        """lw_hip_encode (render/post.rs:141-144): clamp(x, 0, 1)^(1/2.2) as libm's powf."""
        a = np.ascontiguousarray(x, dtype="<f4").reshape(-1)
        out = np.zeros_like(a)
        ms = ctypes.c_double()
        self._check(_native.lib().lw_hip_encode(self.handle, _addr(a), a.size, _addr(out),
                                                ctypes.byref(ms)))
        self.encode_setup_ms = ms.value
        return out

    def casts(self, x: np.ndarray):
        """lw_hip_casts: Rust's f32 `as i32` and `as u32` of each value, as the kernels convert."""
        a = np.ascontiguousarray(x, dtype="<f4").reshape(-1)
        oi = np.zeros(a.size, dtype="<i4")
        ou = np.zeros(a.size, dtype="<u4")
        self._check(_native.lib().lw_hip_casts(self.handle, _addr(a), a.size, _addr(oi), _addr(ou)))
        return oi, ou

    def sky_cache(self, params: np.ndarray) -> np.ndarray:
        """lw_hip_sky_cache: build_sky_cache (integrator.rs:133-148) on the device at the
        resolution params.misc_f[2]."""
        p = np.ascontiguousarray(params, dtype=PARAMS).reshape(())
        n = ctypes.c_size_t()
        L = _native.lib()
        self._check(L.lw_hip_sky_cache(self.handle, _addr(p), None, 0, ctypes.byref(n)))
        out = np.zeros(n.value, dtype="<f4")
        if n.value:
            self._check(L.lw_hip_sky_cache(self.handle, _addr(p), _addr(out), out.size,
                                           ctypes.byref(n)))
        return out

    def blend_chunky_sky(self, params: np.ndarray, weight: float, maps: np.ndarray) -> np.ndarray:
        """lw_hip_blend_chunky_sky (sky.rs:836-851): a blended copy of the two maps."""
        p = np.ascontiguousarray(params, dtype=PARAMS).reshape(())
        out = np.ascontiguousarray(maps, dtype="<f4").copy()
        self._check(_native.lib().lw_hip_blend_chunky_sky(self.handle, _addr(p), weight, _addr(out),
                                                          out.size))
        return out

    def game_sky_maps(self, lens, light, sky_light_factor: float, sky_light_color, atlas,
                      atlas_w: int, stars: np.ndarray):
        """lw_hip_game_sky_maps (sky.rs:906-953): (the two maps as one float32 buffer, the
        lens's then the light's, and a dict of sizes, light, mean, scale, veil and times)."""
        at = np.ascontiguousarray(atlas, dtype="<u4")
        st = np.ascontiguousarray(stars, dtype="<u4")
        col = np.ascontiguousarray(sky_light_color, dtype="<f4")
        out = np.zeros((2048 * 1024 + 360 * 180) * 3, dtype="<f4")
        info = GameSkyC()
        self._check(_native.lib().lw_hip_game_sky_maps(
            self.handle, ctypes.byref(lens), ctypes.byref(light), sky_light_factor, _addr(col),
            _addr(at), at.size, atlas_w, _addr(st), st.size, _addr(out), out.size, ctypes.byref(info)))
        return out, {"sizes": list(info.sizes), "light": list(info.light), "mean": list(info.mean),
                     "scale": info.scale, "veil": list(info.veil), "kernel_ms": info.kernel_ms,
                     "setup_ms": info.setup_ms}

    def downsample(self, px: np.ndarray, w: int, h: int, k: int) -> np.ndarray:
        """lw_hip_downsample (render/post.rs:218-247)."""
        a = np.ascontiguousarray(px, dtype="<u4").reshape(-1)
        n = (w // k) * (h // k) if k > 1 else w * h
        out = np.zeros(n, dtype="<u4")
        self._check(_native.lib().lw_hip_downsample(self.handle, _addr(a), w, h, k, _addr(out)))
        return out

    def render_hq_rgba8(self, scene: HipScene, params: np.ndarray, spp: int,
                        exposure: float = 1.0, op: int = 0) -> np.ndarray:
        """lw_hip_render_hq_rgba8: the frame rendered and tone-mapped on the device."""
        p = np.ascontiguousarray(params, dtype=PARAMS).reshape(())
        out = np.zeros(int(p["frame"][0]) * int(p["frame"][1]), dtype="<u4")
        self._check(_native.lib().lw_hip_render_hq_rgba8(self.handle, scene.handle, _addr(p), spp,
                                                         exposure, op, _addr(out)))
        return out

    def copy_probe(self, nbytes: int) -> float:
        """Times a plain device-to-device copy of nbytes (ms); see lw_hip_copy_probe."""
        ms = ctypes.c_double()
        self._check(_native.lib().lw_hip_copy_probe(self.handle, nbytes, ctypes.byref(ms)))
        return ms.value

    def last_timing(self) -> dict:
        t = TimingC()
        self._check(_native.lib().lw_hip_last_timing(self.handle, ctypes.byref(t)))
        return {"trace_ms": t.trace_ms, "accumulate_ms": t.accumulate_ms, "wall_ms": t.wall_ms,
                "samples": int(t.samples), "launches": int(t.launches)}

    def trace(self, view: BoundView, o, d, tmin=None, tmax=None) -> np.ndarray:
        """trace (trace.rs:137-290) on the device for rays (n, 3); HIT records."""
        o = np.ascontiguousarray(o, dtype="<f4").reshape(-1, 3)
        d = np.ascontiguousarray(d, dtype="<f4").reshape(-1, 3)
        n = o.shape[0]
        tmin = np.ascontiguousarray(np.broadcast_to(0.0 if tmin is None else tmin, (n,)), "<f4")
        tmax = np.ascontiguousarray(np.broadcast_to(np.inf if tmax is None else tmax, (n,)), "<f4")
        out = np.zeros(n, dtype=HIT)
        self._check(_native.lib().lw_hip_trace(self.handle, view.ref, n, _addr(o), _addr(d),
                                               _addr(tmin), _addr(tmax), _addr(out)))
        return out

    def render_samples(self, view: BoundView, px, py, sample) -> np.ndarray:
        """render_sample (integrator.rs:539-709) on the device; (n, 3) float32."""
        px = np.ascontiguousarray(px, dtype="<u4")
        py = np.ascontiguousarray(py, dtype="<u4")
        sample = np.ascontiguousarray(sample, dtype="<u4")
        out = np.zeros((px.size, 3), dtype="<f4")
        self._check(_native.lib().lw_hip_render_samples(self.handle, view.ref, px.size, _addr(px),
                                                        _addr(py), _addr(sample), _addr(out)))
        return out

    def close(self):
        if self.handle:
            _native.lib().lw_hip_close(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class HipFit:
    """An inverse fit resident on the device (lw_hip_fit_*)."""

    def __init__(self, backend: HipBackend, scene: HipScene, handle: int, n_pal: int, n_int: int):
        self.backend, self.scene, self.handle = backend, scene, handle
        self.n_pal, self.n_int = n_pal, n_int

    def step(self) -> dict:
        st = FitStatsC()
        self.backend._check(_native.lib().lw_hip_fit_step(self.backend.handle, self.scene.handle,
                                                          self.handle, ctypes.byref(st)))
        return {"loss": st.loss, "forward_ms": st.forward_ms, "adjoint_ms": st.adjoint_ms,
                "wall_ms": st.wall_ms, "samples": int(st.samples), "step": int(st.step)}

    def params(self):
        pal = np.zeros((self.n_pal, 3), dtype="<f4")
        inten = np.zeros(self.n_int, dtype="<f4")
        grads = np.zeros(3 * self.n_pal + self.n_int + 3, dtype="<f8")
        self.backend._check(_native.lib().lw_hip_fit_params(
            self.backend.handle, self.handle, _addr(pal) if pal.size else None,
            _addr(inten) if inten.size else None, _addr(grads)))
        return pal, inten, grads

    def free(self):
        if self.handle:
            _native.lib().lw_hip_fit_free(self.backend.handle, self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
