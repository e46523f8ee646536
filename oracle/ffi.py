"""ORACLE -- test infrastructure only.

ctypes binding of the C++ restatement of the reference kernel crate
(oracle/lw_oracle_impl.hpp). Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg import this module, and only as the checker or the CPU baseline.
"""

from __future__ import annotations

import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

ORACLE_DIR = Path(__file__).resolve().parent
LIB_PATH = ORACLE_DIR / "_build" / "liblw_oracle.so"

_lib = None


def build(force: bool = False) -> Path:
    """Compiles the oracle with its Makefile (g++, no FMA contraction)."""
    if force or not LIB_PATH.exists():
        subprocess.run(["make", "-s", "-C", str(ORACLE_DIR)], check=True)
    return LIB_PATH


class Counters(ctypes.Structure):
    _fields_ = [("bytes", ctypes.c_uint64), ("dda_steps", ctypes.c_uint64),
                ("quads_tested", ctypes.c_uint64), ("traces", ctypes.c_uint64),
                ("samples", ctypes.c_uint64), ("sky_lookups", ctypes.c_uint64),
                ("see_through_drawn", ctypes.c_uint64), ("see_through_forced", ctypes.c_uint64),
                ("translucent_passes", ctypes.c_uint64), ("alpha_zero_rejects", ctypes.c_uint64),
                ("shadow_water_exits", ctypes.c_uint64), ("shadow_translucent", ctypes.c_uint64),
                ("water_interfaces", ctypes.c_uint64), ("pad", ctypes.c_uint64)]


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = ctypes.CDLL(str(LIB_PATH))
        vp, sz, u32, f32, i32 = (ctypes.c_void_p, ctypes.c_size_t, ctypes.c_uint32,
                                 ctypes.c_float, ctypes.c_int)
        L.lwo_last_error.restype = ctypes.c_char_p
        L.lwo_sizeof.argtypes = [i32]
        L.lwo_sizeof.restype = u32
        L.lwo_pcg_hash.argtypes = [u32]
        L.lwo_pcg_hash.restype = u32
        L.lwo_rng_new.argtypes = [u32, u32, u32]
        L.lwo_rng_new.restype = u32
        L.lwo_rng_draws.argtypes = [u32, u32, vp]
        L.lwo_rng_draws.restype = u32
        for name in ("lwo_acos", "lwo_ln", "lwo_gamma_to_linear", "lwo_expf"):
            getattr(L, name).argtypes = [vp, sz, vp]
            getattr(L, name).restype = None
        L.lwo_sin_cos_turns.argtypes = [vp, sz, vp, vp]
        L.lwo_sin_cos_turns.restype = None
        L.lwo_atan2.argtypes = [vp, vp, sz, vp]
        L.lwo_atan2.restype = None
        L.lwo_onb.argtypes = [vp, vp, vp]
        L.lwo_onb.restype = None
        L.lwo_trace.argtypes = [vp, sz, vp, vp, vp, vp, vp]
        L.lwo_trace.restype = i32
        L.lwo_is_see_through.argtypes = [vp, u32]
        L.lwo_is_see_through.restype = i32
        L.lwo_see_through_hit.argtypes = [vp, vp, vp, f32, vp]
        L.lwo_see_through_hit.restype = i32
        L.lwo_render_samples.argtypes = [vp, sz, vp, vp, vp, vp, ctypes.c_uint]
        L.lwo_render_samples.restype = i32
        L.lwo_render_hq.argtypes = [vp, u32, ctypes.c_uint, vp]
        L.lwo_render_hq.restype = i32
        L.lwo_render_hq_counted.argtypes = [vp, u32, ctypes.c_uint, vp, vp, i32]
        L.lwo_render_hq_counted.restype = i32
        L.lwo_octant_boxes.argtypes = [vp, vp]
        L.lwo_octant_boxes.restype = i32
        L.lwo_render_diff.argtypes = [vp, u32, vp, ctypes.c_uint, vp]
        L.lwo_render_diff.restype = i32
        L.lwo_adjoint.argtypes = [vp, u32, vp, vp, ctypes.c_uint, vp]
        L.lwo_adjoint.restype = i32
        L.lwo_render_diff_dot.argtypes = [vp, u32, vp, vp, ctypes.c_uint, vp]
        L.lwo_render_diff_dot.restype = i32
        L.lwo_tonemap.argtypes = [vp, sz, f32, u32, vp]
        L.lwo_tonemap.restype = None
        L.lwo_tonemap_mix.argtypes = [vp, sz, f32, u32, u32, f32, vp]
        L.lwo_tonemap_mix.restype = None
        L.lwo_downsample.argtypes = [vp, u32, u32, u32, vp]
        L.lwo_downsample.restype = None
        L.lwo_hable.argtypes = [u32, vp, ctypes.POINTER(ctypes.c_double), ctypes.c_double,
                                ctypes.POINTER(ctypes.c_double)]
        L.lwo_hable.restype = None
        L.lwo_encode.argtypes = [vp, sz, vp]
        L.lwo_encode.restype = None
        L.lwo_pfm.argtypes = [vp, u32, u32, vp, sz]
        L.lwo_pfm.restype = sz
        L.lwo_check_encode_thresholds.argtypes = [vp, ctypes.c_uint, ctypes.POINTER(u32)]
        L.lwo_check_encode_thresholds.restype = ctypes.c_uint64
        L.lwo_powf_band.argtypes = [ctypes.c_float, ctypes.c_uint, vp, vp, vp, ctypes.c_uint64]
        L.lwo_powf_band.restype = None
        L.lwo_java_random.argtypes = [ctypes.c_int64, i32, sz, vp]
        L.lwo_java_random.restype = None
        L.lwo_stars.argtypes = [vp, sz]
        L.lwo_stars.restype = ctypes.c_long
        L.lwo_star_table.argtypes = [vp, sz]
        L.lwo_star_table.restype = ctypes.c_long
        L.lwo_quick_sky_new.argtypes = [vp, u32, u32, vp, vp]
        L.lwo_quick_sky_new.restype = None
        L.lwo_celestial_direction.argtypes = [ctypes.c_float, vp]
        L.lwo_celestial_direction.restype = None
        L.lwo_lightmap_sky.argtypes = [ctypes.c_float, vp, ctypes.c_float, vp]
        L.lwo_lightmap_sky.restype = None
        L.lwo_quick_background.argtypes = [vp, vp, sz, vp, u32, vp]
        L.lwo_quick_background.restype = None
        L.lwo_game_sky_maps.argtypes = [vp, vp, ctypes.c_float, vp, vp, u32, ctypes.c_uint, vp, vp]
        L.lwo_game_sky_maps.restype = None
        L.lwo_blend_chunky_sky.argtypes = [vp, ctypes.c_float, vp, vp]
        L.lwo_blend_chunky_sky.restype = i32
        L.lwo_log2f_band.argtypes = [ctypes.c_uint, vp, vp, vp, ctypes.c_uint64]
        L.lwo_log2f_band.restype = None
        L.lwo_rasterize.argtypes = [vp, sz, vp, vp, vp, vp, vp, u32, u32, u32, u32, vp, vp, ctypes.c_uint, vp]
        L.lwo_rasterize.restype = i32
        L.lwo_atlas_sample.argtypes = [vp, vp, vp, u32, u32, i32, vp, sz, vp]
        L.lwo_atlas_sample.restype = None
        L.lwo_shadow.argtypes = [vp, sz, vp, vp, vp, vp]
        L.lwo_shadow.restype = i32
        L.lwo_sun_texture.argtypes = [vp, sz, vp, vp]
        L.lwo_sun_texture.restype = i32
        L.lwo64_render_hq.argtypes = [vp, u32, ctypes.c_uint, vp]
        L.lwo64_render_hq.restype = i32
        L.lwo64_render_samples.argtypes = [vp, sz, vp, vp, vp, vp]
        L.lwo64_render_samples.restype = i32
        L.lwo64_render_rows.argtypes = [vp, u32, ctypes.c_uint, vp, sz, vp]
        L.lwo64_render_rows.restype = i32
        L.lwo_build_layout.argtypes = [vp, vp, vp, vp, sz, vp, sz, vp, sz, u32, vp, vp, vp]
        L.lwo_build_layout.restype = i32
        L.lwo_build_fetch.argtypes = [vp, vp, vp]
        L.lwo_build_fetch.restype = None
        L.lwo_sky_tangent.argtypes = [vp, vp, i32, vp, sz, vp]
        L.lwo_sky_tangent.restype = i32
        L.lwo_preetham_radiance.argtypes = [vp, vp, sz, vp]
        L.lwo_preetham_radiance.restype = i32
        L.lwo_rgb_to_hsl.argtypes = [vp, sz, vp]
        L.lwo_rgb_to_hsl.restype = None
        L.lwo_rgb_from_hsl.argtypes = [vp, sz, vp]
        L.lwo_rgb_from_hsl.restype = None
        L.lwo_build_sky_cache.argtypes = [vp, u32, vp, sz]
        L.lwo_build_sky_cache.restype = i32
        L.lwo_sky.argtypes = [vp, i32, vp, sz, vp]
        L.lwo_sky.restype = i32
        L.lwo_emitter_cell.argtypes = [vp, vp, vp]
        L.lwo_emitter_cell.restype = i32
        L.lwo_sample_emitters.argtypes = [vp, vp, vp, vp, u32, sz, vp]
        L.lwo_sample_emitters.restype = i32
        L.lwo_water_interface.argtypes = [vp, vp, vp, vp, i32, i32, f32, u32, vp, vp, vp]
        L.lwo_water_interface.restype = u32
        _lib = L
    return _lib


def _check(rc: int):
    if rc != 0:
        raise RuntimeError("oracle: " + lib().lwo_last_error().decode())


def _f32(a):
    return np.ascontiguousarray(a, dtype="<f4")


def _addr(a):
    return ctypes.c_void_p(a.ctypes.data)


def _params_addr(view):
    return ctypes.c_void_p(view.params.ctypes.data)


# ---- math (crates/lumenwright-kernel/src/math.rs) ----

def pcg_hash(v: int) -> int:
    return lib().lwo_pcg_hash(v & 0xFFFFFFFF)


def rng_new(pixel: int, sample: int, seed: int) -> int:
    return lib().lwo_rng_new(pixel, sample, seed)


def rng_draws(state: int, n: int):
    out = np.zeros(n, dtype="<f4")
    st = lib().lwo_rng_draws(state, n, _addr(out))
    return out, st


def _unary(name, x):
    x = _f32(x)
    out = np.empty_like(x)
    getattr(lib(), name)(_addr(x), x.size, _addr(out))
    return out


def acos(x):
    return _unary("lwo_acos", x)


def ln(x):
    return _unary("lwo_ln", x)


def gamma_to_linear(x):
    return _unary("lwo_gamma_to_linear", x)


def expf(x):
    return _unary("lwo_expf", x)


def sin_cos_turns(u):
    u = _f32(u)
    s, c = np.empty_like(u), np.empty_like(u)
    lib().lwo_sin_cos_turns(_addr(u), u.size, _addr(s), _addr(c))
    return s, c


def atan2(y, x):
    y, x = _f32(y), _f32(x)
    out = np.empty_like(y)
    lib().lwo_atan2(_addr(y), _addr(x), y.size, _addr(out))
    return out


def onb(n):
    n = _f32(n)
    t, b = np.empty(3, "<f4"), np.empty(3, "<f4")
    lib().lwo_onb(_addr(n), _addr(t), _addr(b))
    return t, b


# ---- trace.rs ----

def trace(view, o, d, tmin=None, tmax=None):
    """trace (trace.rs:137-290) for rays o, d of shape (n, 3); returns HIT records."""
    from lumenwright_amd.abi import HIT
    o = _f32(o).reshape(-1, 3)
    d = _f32(d).reshape(-1, 3)
    n = o.shape[0]
    tmin = _f32(np.zeros(n) if tmin is None else np.broadcast_to(tmin, (n,)))
    tmax = _f32(np.full(n, np.inf) if tmax is None else np.broadcast_to(tmax, (n,)))
    out = np.zeros(n, dtype=HIT)
    _check(lib().lwo_trace(view.ref, n, _addr(o), _addr(d), _addr(tmin), _addr(tmax), _addr(out)))
    return out


def shadow(view, o, d, tmax):
    """shadow (integrator.rs:316-346): the transmittance along rays o, d (n, 3) up to tmax."""
    o = _f32(o).reshape(-1, 3)
    d = _f32(d).reshape(-1, 3)
    tm = _f32(np.broadcast_to(tmax, (o.shape[0],)))
    out = np.zeros_like(o)
    _check(lib().lwo_shadow(view.ref, o.shape[0], _addr(o), _addr(d), _addr(tm), _addr(out)))
    return out


def sun_texture(view, d):
    """sun_texture (integrator.rs:252-277) along directions d (n, 3)."""
    d = _f32(d).reshape(-1, 3)
    out = np.zeros_like(d)
    _check(lib().lwo_sun_texture(view.ref, d.shape[0], _addr(d), _addr(out)))
    return out


def is_see_through(view, qi):
    return bool(lib().lwo_is_see_through(view.ref, qi))


def see_through_hit(view, o, d, tmin):
    from lumenwright_amd.abi import HIT
    out = np.zeros((), dtype=HIT)
    _check(lib().lwo_see_through_hit(view.ref, _addr(_f32(o)), _addr(_f32(d)), float(tmin),
                                     _addr(out)))
    return out if out["hit"] else None


# ---- integrator.rs ----

def render_samples(view, px, py, sample, threads: int = 0):
    """render_sample (integrator.rs:539-709) for each (px, py, sample); (n, 3) f32."""
    px = np.ascontiguousarray(px, dtype="<u4")
    py = np.ascontiguousarray(py, dtype="<u4")
    sample = np.ascontiguousarray(sample, dtype="<u4")
    n = px.size
    out = np.zeros((n, 3), dtype="<f4")
    threads = threads or os.cpu_count() or 1
    _check(lib().lwo_render_samples(view.ref, n, _addr(px), _addr(py), _addr(sample), _addr(out),
                                    threads))
    return out


def render_diff(view, spp: int, desc, threads: int = 0):
    """The CPU reference of lw_hip_render_diff: (h, w, 3) f32 mean radiance under the
    parameters of `desc` (a lumenwright_amd.abi.DiffDescC)."""
    w, h = int(view.params["frame"][0]), int(view.params["frame"][1])
    out = np.zeros((h, w, 3), dtype="<f4")
    threads = threads or os.cpu_count() or 1
    _check(lib().lwo_render_diff(view.ref, spp, ctypes.byref(desc), threads, _addr(out)))
    return out


def render_diff_dot(view, spp: int, desc, weights, threads: int = 0) -> float:
    """sum(weights * image) of render_diff, the image kept in f64."""
    g = np.ascontiguousarray(weights, dtype="<f4").reshape(-1)
    out = ctypes.c_double()
    threads = threads or os.cpu_count() or 1
    _check(lib().lwo_render_diff_dot(view.ref, spp, ctypes.byref(desc), _addr(g), threads,
                                     ctypes.byref(out)))
    return out.value


def adjoint(view, spp: int, desc, n_grads: int, dldi, threads: int = 0):
    """The CPU reference gradient (path replay) of sum(dldi * image): (n_grads,) f64."""
    g = np.ascontiguousarray(dldi, dtype="<f4").reshape(-1)
    out = np.zeros(n_grads, dtype="<f8")
    threads = threads or os.cpu_count() or 1
    _check(lib().lwo_adjoint(view.ref, spp, ctypes.byref(desc), _addr(g), threads, _addr(out)))
    return out


def sky_tangent(view, desc, k: int, d):
    """d sky_radiance(d) / d(sky parameter k) through the desc's derivative tables."""
    d = _f32(d).reshape(-1, 3)
    out = np.zeros_like(d)
    _check(lib().lwo_sky_tangent(view.ref, ctypes.byref(desc), k, _addr(d), d.shape[0], _addr(out)))
    return out


def render_hq(view, spp: int, threads: int = 0):
    """CpuBackend::render_hq (render/cpu/mod.rs:39-87); (h, w, 3) f32 mean radiance."""
    w, h = int(view.params["frame"][0]), int(view.params["frame"][1])
    out = np.zeros((h, w, 3), dtype="<f4")
    threads = threads or os.cpu_count() or 1
    _check(lib().lwo_render_hq(view.ref, spp, threads, _addr(out)))
    return out


def render_hq_f64(view, spp: int, threads: int = 0):
    """CpuBackend::render_hq with every operation in f64 and libm's double functions
    (lw_oracle_f64.cpp): (h, w, 3) float64, as render_hq's shape."""
    w, h = int(view.params["frame"][0]), int(view.params["frame"][1])
    out = np.zeros((h, w, 3), dtype=np.float64)
    rc = lib().lwo64_render_hq(view.ref, spp, threads or os.cpu_count() or 1, _addr(out))
    if rc != 0:
        raise RuntimeError("oracle: f64 render failed")
    return out


def render_rows_f64(view, spp: int, rows, threads: int = 0):
    """render_hq_f64 on the listed rows only: (len(rows), w, 3) float64."""
    w = int(view.params["frame"][0])
    r = np.ascontiguousarray(rows, dtype="<u4")
    out = np.zeros((r.size, w, 3), dtype=np.float64)
    if lib().lwo64_render_rows(view.ref, spp, threads or os.cpu_count() or 1, _addr(r), r.size,
                               _addr(out)) != 0:
        raise RuntimeError("oracle: f64 row render failed")
    return out


def render_samples_f64(view, px, py, sample):
    """render_sample in f64 for (px, py, sample) triples: (n, 3) float64."""
    px, py, sm = (np.ascontiguousarray(a, dtype="<u4") for a in (px, py, sample))
    out = np.zeros((px.size, 3), dtype=np.float64)
    if lib().lwo64_render_samples(view.ref, px.size, _addr(px), _addr(py), _addr(sm), _addr(out)) != 0:
        raise RuntimeError("oracle: f64 samples failed")
    return out


def render_hq_counted(view, spp: int, threads: int = 0, device_walk: bool = True):
    """render_hq with the algorithmic-byte and branch counters; returns (image, Counters).
    device_walk: count the walk the device runs (stopped once every brick left toward the
    ray's octant is empty); False: the reference's whole walk."""
    w, h = int(view.params["frame"][0]), int(view.params["frame"][1])
    out = np.zeros((h, w, 3), dtype="<f4")
    c = Counters()
    threads = threads or os.cpu_count() or 1
    _check(lib().lwo_render_hq_counted(view.ref, spp, threads, _addr(out), ctypes.byref(c),
                                       1 if device_walk else 0))
    return out, c


def octant_boxes(view):
    """Per brick, the octant bits the counted walk (and the device's walk) stops by: bit k
    (k = [+x] | [+y] << 1 | [+z] << 2) set when the box toward octant k holds a brick that is
    not EMPTY. Bricks are numbered (y * dz + z) * dx + x."""
    dims = [int(x) for x in view.params["grid_dims"][:3]]
    out = np.zeros(dims[0] * dims[1] * dims[2], dtype="u1")
    _check(lib().lwo_octant_boxes(view.ref, _addr(out)))
    return out


def preetham_radiance(view, d):
    d = _f32(d).reshape(-1, 3)
    out = np.zeros_like(d)
    _check(lib().lwo_preetham_radiance(view.ref, _addr(d), d.shape[0], _addr(out)))
    return out


def rgb_to_hsl(rgb):
    rgb = _f32(rgb).reshape(-1, 3)
    out = np.zeros_like(rgb)
    lib().lwo_rgb_to_hsl(_addr(rgb), rgb.shape[0], _addr(out))
    return out


def rgb_from_hsl(hsl):
    hsl = _f32(hsl).reshape(-1, 3)
    out = np.zeros_like(hsl)
    lib().lwo_rgb_from_hsl(_addr(hsl), hsl.shape[0], _addr(out))
    return out


def build_sky_cache(view, n: int):
    out = np.zeros((n + 1) * (n + 1) * 3, dtype="<f4")
    _check(lib().lwo_build_sky_cache(view.ref, n, _addr(out), out.size))
    return out


def _sky(view, which, d):
    d = _f32(d).reshape(-1, 3)
    out = np.zeros_like(d)
    _check(lib().lwo_sky(view.ref, which, _addr(d), d.shape[0], _addr(out)))
    return out


def sky_radiance(view, d):
    return _sky(view, 0, d)


def sky_lens(view, d):
    return _sky(view, 1, d)


def sky_light(view, d):
    return _sky(view, 2, d)


def sun_texture(view, d):
    return _sky(view, 3, d)


def emitter_cell(view, p):
    out = np.zeros(2, dtype="<u4")
    _check(lib().lwo_emitter_cell(view.ref, _addr(_f32(p)), _addr(out)))
    return int(out[0]), int(out[1])


def sample_emitters(view, pos, origin, n, rng_state: int, draws: int):
    out = np.zeros((draws, 3), dtype="<f4")
    _check(lib().lwo_sample_emitters(view.ref, _addr(_f32(pos)), _addr(_f32(origin)),
                                     _addr(_f32(n)), rng_state, draws, _addr(out)))
    return out


def water_interface(params, pos, d, n, front, in_water, eps, rng_state):
    d_out, o_out = np.zeros(3, "<f4"), np.zeros(3, "<f4")
    w = ctypes.c_int(0)
    st = lib().lwo_water_interface(ctypes.c_void_p(params.ctypes.data), _addr(_f32(pos)),
                                   _addr(_f32(d)), _addr(_f32(n)), int(front), int(in_water),
                                   float(eps), rng_state, _addr(d_out), _addr(o_out),
                                   ctypes.byref(w))
    return d_out, o_out, bool(w.value), st


# ---- post-processing (crates/lumenwright/src/render/post.rs) ----

def tonemap(hdr, exposure: float, op: int):
    h = _f32(hdr).reshape(-1, 3)
    out = np.zeros(h.shape[0], dtype="<u4")
    lib().lwo_tonemap(_addr(h), h.shape[0], exposure, op, _addr(out))
    return out


def tonemap_mix(hdr, exposure: float, a: int, b: int, weight: float):
    h = _f32(hdr).reshape(-1, 3)
    out = np.zeros(h.shape[0], dtype="<u4")
    lib().lwo_tonemap_mix(_addr(h), h.shape[0], exposure, a, b, weight, _addr(out))
    return out


def downsample(px, w: int, h: int, k: int):
    a = np.ascontiguousarray(px, dtype="<u4").reshape(-1)
    n = (w // k) * (h // k) if k > 1 else w * h
    out = np.zeros(n, dtype="<u4")
    lib().lwo_downsample(_addr(a), w, h, k, _addr(out))
    return out


def hable(op: int, x: float = 0.18):
    """(a, b, c, d, e, f, w, white_scale), mid_grey_gain(0.18), before_gamma(x)."""
    out = np.zeros(8, dtype="<f4")
    g, bg = ctypes.c_double(), ctypes.c_double()
    lib().lwo_hable(op, _addr(out), ctypes.byref(g), x, ctypes.byref(bg))
    return out, g.value, bg.value


def encode(v):
    v = _f32(v)
    out = np.empty_like(v)
    lib().lwo_encode(_addr(v), v.size, _addr(out))
    return out


def pfm(hdr, w: int, h: int) -> bytes:
    a = _f32(hdr).reshape(-1)
    n = lib().lwo_pfm(_addr(a), w, h, None, 0)
    buf = np.zeros(n, dtype=np.uint8)
    lib().lwo_pfm(_addr(a), w, h, _addr(buf), n)
    return buf.tobytes()


def check_encode_thresholds(thr, threads: int = 0):
    """Mismatches of the threshold byte against post.rs's byte over every f32 in [0, 1]."""
    t = _f32(thr)
    first = ctypes.c_uint32()
    n = lib().lwo_check_encode_thresholds(_addr(t), threads or os.cpu_count() or 1,
                                          ctypes.byref(first))
    return n, first.value


def powf_band(threads: int = 0, inputs: bool = False, exponent: float = 1.0 / 2.2):
    """libm's powf(x, y) (y the f32 of `exponent`, 1/2.2 by default) against the rounded f64
    power over every f32 in [0, 1]:
    (deviations, deviations outside the 2^-32 band, inputs in the band, deviations that are not
    neighbours, largest relative midpoint distance of a deviation[, the band inputs])."""
    st = np.zeros(4, dtype=np.uint64)
    mr = ctypes.c_double()
    band = np.zeros(1 << 24, dtype="<u4") if inputs else None
    y = float(np.float32(1.0) / np.float32(2.2)) if exponent == 1.0 / 2.2 else float(np.float32(exponent))
    lib().lwo_powf_band(y, threads or os.cpu_count() or 1, _addr(st),
                        ctypes.byref(mr), _addr(band) if inputs else None,
                        band.size if inputs else 0)
    out = (int(st[0]), int(st[1]), int(st[2]), int(st[3]), mr.value)
    return out + (band[: min(int(st[2]), band.size)],) if inputs else out


def build_layout(vol):
    """build_hq's layout of a VoxelVolume (lumenwright_amd.abi): (grid, emitters, emitter grid,
    dict of offsets, see_through, emitter_count and the brick counts)."""
    from lumenwright_amd.abi import EMITTER, MODEL, QUAD
    dims = np.array(vol.dims, dtype="<u4")
    vox = np.ascontiguousarray(vol.voxels, dtype="<u2").reshape(-1)
    occ = None if vol.brick_occludes is None else np.ascontiguousarray(vol.brick_occludes, np.uint8)
    models = np.ascontiguousarray(vol.models, dtype=MODEL)
    flags = np.ascontiguousarray(vol.model_flags, dtype=np.uint8)
    ent = np.ascontiguousarray(vol.entity_quads, dtype=QUAD)
    sizes = np.zeros(3, dtype=np.uint64)
    meta = np.zeros(8, dtype="<u4")
    stats = np.zeros(4, dtype=np.uint64)
    rc = lib().lwo_build_layout(_addr(dims), _addr(vox), _addr(occ) if occ is not None else None,
                                _addr(models), models.size, _addr(flags), len(vol.quads),
                                _addr(ent) if ent.size else None, ent.size, vol.see_through_count,
                                _addr(sizes), _addr(meta), _addr(stats))
    if rc != 0:
        raise RuntimeError(lib().lwo_last_error().decode())
    grid = np.zeros(int(sizes[0]), dtype="<u4")
    em = np.zeros(int(sizes[1]), dtype=EMITTER)
    eg = np.zeros(int(sizes[2]), dtype="<u4")
    lib().lwo_build_fetch(_addr(grid), _addr(em), _addr(eg))
    info = {"offsets": tuple(int(x) for x in meta[:5]), "see_through": (int(meta[5]), int(meta[6])),
            "emitter_count": int(meta[7]), "bricks": int(stats[0]), "uniform_bricks": int(stats[1]),
            "buried_bricks": int(stats[2]), "overlay_bricks": int(stats[3])}
    return grid, em, eg, info


def blend_chunky_sky(view, weight: float, sizes, maps):
    """blend_chunky_sky (sky.rs:836-851) of a copy of `maps` with the view's sky cache."""
    out = np.ascontiguousarray(maps, dtype="<f4").copy()
    sz = np.ascontiguousarray(sizes, dtype="<f4")
    _check(lib().lwo_blend_chunky_sky(view.ref, weight, _addr(sz), _addr(out)))
    return out


# ---- the game's sky (render/sky.rs) ----

def java_random(seed: int, bits: int, n: int):
    out = np.zeros(n, dtype=np.int32)
    lib().lwo_java_random(seed, bits, n, _addr(out))
    return out


def stars():
    n = lib().lwo_stars(None, 0)
    out = np.zeros((n, 9), dtype="<f4")
    lib().lwo_stars(_addr(out), out.size)
    return out


def star_table():
    n = lib().lwo_star_table(None, 0)
    out = np.zeros(n, dtype="<u4")
    lib().lwo_star_table(_addr(out), n)
    return out


def quick_sky_new(cam, w: int, h: int, desc):
    """QuickSky::new; cam is a CameraRaysC, desc a SkyDescriptionC (lumenwright_amd.frames)."""
    from lumenwright_amd.frames import QuickSkyC
    c = np.array(list(cam.forward) + list(cam.right) + list(cam.up) + [cam.tan_half_fov, cam.aspect],
                 dtype="<f4")
    q = QuickSkyC()
    lib().lwo_quick_sky_new(_addr(c), w, h, ctypes.byref(desc), ctypes.byref(q))
    return q


def celestial_direction(deg: float):
    out = np.zeros(3, dtype="<f4")
    lib().lwo_celestial_direction(deg, _addr(out))
    return out


def lightmap_sky(factor: float, colour, brightness: float = 0.5):
    out = np.zeros(3, dtype="<f4")
    lib().lwo_lightmap_sky(factor, _addr(_f32(colour)), brightness, _addr(out))
    return out


def quick_background(q, d, atlas, atlas_w: int):
    d = _f32(d).reshape(-1, 3)
    out = np.zeros_like(d)
    at = np.ascontiguousarray(atlas, dtype="<u4")
    lib().lwo_quick_background(ctypes.byref(q), _addr(d), d.shape[0], _addr(at), atlas_w, _addr(out))
    return out


def game_sky_maps(lens, light, factor: float, colour, atlas, atlas_w: int, threads: int = 0):
    """game_sky_maps, the texels on `threads` threads (all cores by default): (the two maps,
    dict of light, mean, scale, veil)."""
    out = np.zeros((2048 * 1024 + 360 * 180) * 3, dtype="<f4")
    info = np.zeros(10, dtype="<f4")
    at = np.ascontiguousarray(atlas, dtype="<u4")
    lib().lwo_game_sky_maps(ctypes.byref(lens), ctypes.byref(light), factor, _addr(_f32(colour)),
                            _addr(at), atlas_w, threads or os.cpu_count() or 1, _addr(out), _addr(info))
    return out, {"light": list(info[0:3]), "mean": list(info[3:6]), "scale": float(info[6]),
                 "veil": list(info[7:10])}


def log2f_band(threads: int = 0):
    """libm's log2f against the rounded f64 log2 over every f32 in [1, 65536], as powf_band."""
    st = np.zeros(4, dtype=np.uint64)
    mr = ctypes.c_double()
    lib().lwo_log2f_band(threads or os.cpu_count() or 1, _addr(st), ctypes.byref(mr), None, 0)
    return (int(st[0]), int(st[1]), int(st[2]), int(st[3]), mr.value)


def rasterize(frame, threads: int = 1):
    """rasterize (raster.rs:168-372) of a lumenwright_amd.abi.QuickFrame with the client's
    stars, its tiles on `threads` threads: width * height packed RGBA8 words."""
    from lumenwright_amd.abi import QUICK_VERTEX
    m = frame.mesh
    v = np.ascontiguousarray(m.vertices, dtype=QUICK_VERTEX)
    lists = [np.ascontiguousarray(a, dtype="<u4") for a in (m.opaque, m.translucent, m.see_through)]
    ptrs = (ctypes.c_void_p * 3)(*[a.ctypes.data for a in lists])
    lens = (ctypes.c_size_t * 3)(*[a.size for a in lists])
    atlas = np.ascontiguousarray(frame.atlas_pixels, dtype="<u4")
    mips = np.ascontiguousarray(frame.atlas_mips, dtype="<u4")
    size = np.array(frame.atlas_size, dtype="<u4")
    vp = _f32(np.asarray(frame.view_proj).reshape(16))
    out = np.zeros(int(frame.width) * int(frame.height), dtype="<u4")
    _check(lib().lwo_rasterize(_addr(v), v.size, ptrs, lens, _addr(atlas), _addr(size),
                               _addr(mips) if mips.size else None, frame.mip_levels, frame.mip_rows,
                               frame.width, frame.height, _addr(vp), ctypes.byref(frame.sky),
                               threads, _addr(out)))
    return out


def atlas_sample(level0, mips, size, levels: int, mip_rows: int, q, mode: int = 0):
    """AtlasView (texture.rs:46-145) at points q (n, 6) = (u, v, du.x, du.y, dv.x, dv.y):
    mode 0 sample, 1 sample_block, 2 sample_grad. Returns (n, 4) RGBA."""
    q = _f32(q).reshape(-1, 6)
    l0 = np.ascontiguousarray(level0, dtype="<u4")
    mp = np.ascontiguousarray(mips, dtype="<u4")
    sz = np.array(size, dtype="<u4")
    out = np.zeros((q.shape[0], 4), dtype="<f4")
    lib().lwo_atlas_sample(_addr(l0), _addr(mp) if mp.size else None, _addr(sz), levels, mip_rows,
                           mode, _addr(q), q.shape[0], _addr(out))
    return out
