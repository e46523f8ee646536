"""ctypes binding of lumenwright_amd/_lib/liblumenwright_hip.so (include/lw/abi.h, host.h).

The library is built in-tree by ``python -m lumenwright_amd.build``. Loading fails loudly
when it is missing: there is no CPU fallback behind this package.
"""

from __future__ import annotations

import ctypes
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "_lib" / "liblumenwright_hip.so"

_lib = None

vp, sz, u32, i32, f32, f64 = (ctypes.c_void_p, ctypes.c_size_t, ctypes.c_uint32, ctypes.c_int,
                              ctypes.c_float, ctypes.c_double)

PROGRESS_FN = ctypes.CFUNCTYPE(None, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_void_p)

# Every C-ABI symbol of include/lw/abi.h and include/lw/host.h: (restype, argtypes).
SIGNATURES = {
    # abi.h
    "lw_hip_probe": (i32, [ctypes.c_char_p, sz]),
    "lw_hip_open": (i32, [i32, ctypes.POINTER(vp)]),
    "lw_hip_name": (ctypes.c_char_p, [vp]),
    "lw_hip_render_hq": (i32, [vp, vp, u32, PROGRESS_FN, vp, vp]),
    "lw_hip_render_quick": (i32, [vp, vp, vp]),
    "lw_hip_last_error": (ctypes.c_char_p, [vp]),
    "lw_last_error": (ctypes.c_char_p, []),
    "lw_hip_close": (None, [vp]),
    "lw_hip_trace": (i32, [vp, vp, sz, vp, vp, vp, vp, vp]),
    "lw_hip_render_samples": (i32, [vp, vp, sz, vp, vp, vp, vp]),
    "lw_hip_scene_upload": (i32, [vp, vp, ctypes.POINTER(vp)]),
    "lw_hip_render_hq_scene": (i32, [vp, vp, vp, u32, PROGRESS_FN, vp, vp]),
    "lw_hip_render_hq_scene_noalloc": (i32, [vp, vp, vp, u32, PROGRESS_FN, vp, vp]),
    "lw_hip_count_traces": (i32, [vp, vp, vp, u32, vp]),
    "lw_hip_scene_free": (None, [vp, vp]),
    "lw_hip_scene_build": (i32, [vp, vp, vp, vp, ctypes.POINTER(vp)]),
    "lw_hip_scene_read": (i32, [vp, vp, u32, vp, sz, ctypes.POINTER(sz)]),
    "lw_hip_last_timing": (i32, [vp, vp]),
    "lw_hip_render_hq_batch": (i32, [vp, vp, vp, sz, PROGRESS_FN, vp, vp]),
    "lw_hip_copy_probe": (i32, [vp, sz, ctypes.POINTER(f64)]),
    "lw_hip_render_diff": (i32, [vp, vp, vp, u32, vp, vp, vp]),
    "lw_hip_adjoint": (i32, [vp, vp, vp, u32, vp, vp, vp]),
    "lw_hip_render_diff_batch": (i32, [vp, vp, vp, sz, vp, vp]),
    "lw_hip_adjoint_batch": (i32, [vp, vp, vp, sz, vp, vp, vp]),
    "lw_hip_fit_new": (i32, [vp, vp, vp, sz, vp, vp, vp, ctypes.POINTER(vp)]),
    "lw_hip_fit_step": (i32, [vp, vp, vp, vp]),
    "lw_hip_fit_params": (i32, [vp, vp, vp, vp, vp]),
    "lw_hip_fit_free": (None, [vp, vp]),
    "lw_hip_autotune": (i32, [vp, vp, vp, u32, vp]),
    "lw_hip_tonemap": (i32, [vp, vp, sz, f32, u32, vp]),
    "lw_hip_tonemap_mix": (i32, [vp, vp, sz, f32, u32, u32, f32, vp]),
    "lw_hip_encode": (i32, [vp, vp, sz, vp, vp]),
    "lw_hip_casts": (i32, [vp, vp, sz, vp, vp]),
    "lw_hip_sky_cache": (i32, [vp, vp, vp, sz, ctypes.POINTER(sz)]),
    "lw_hip_blend_chunky_sky": (i32, [vp, vp, f32, vp, sz]),
    "lw_hip_game_sky_maps": (i32, [vp, vp, vp, f32, vp, vp, sz, u32, vp, sz, vp, sz, vp]),
    "lw_hip_downsample": (i32, [vp, vp, u32, u32, u32, vp]),
    "lw_hip_render_hq_rgba8": (i32, [vp, vp, vp, u32, f32, u32, vp]),
    # host.h
    "lw_celestial_angle": (f32, [ctypes.c_int64]),
    "lw_sun_direction_from_angle": (None, [f32, vp]),
    "lw_sun_direction": (None, [f32, f32, vp]),
    "lw_twilight": (f32, [f32]),
    "lw_hq_sky_new": (None, [vp, f32, f32, vp]),
    "lw_preetham_new": (None, [vp, f64, vp]),
    "lw_sun_texture_axes": (None, [vp, vp, vp]),
    "lw_sky_cache": (ctypes.c_long, [vp, vp, sz]),
    "lw_gamma_lut": (None, [vp]),
    "lw_map_trig": (None, [u32, u32, vp, vp]),
    "lw_quick_sky_new": (None, [vp, u32, u32, vp, vp]),
    "lw_celestial_direction": (None, [f32, vp]),
    "lw_star_table": (ctypes.c_long, [vp, sz]),
    "lw_disc_sector_trig": (None, [vp, vp, vp, vp]),
    "lw_lightmap_sky": (None, [f32, vp, f32, vp]),
    "lw_expf_tabulated": (None, [vp, sz, vp]),
    "lw_camera_basis": (None, [f32, f32, vp, vp, vp]),
    "lw_camera_rays_new": (None, [f32, f32, f32, u32, u32, vp]),
    "lw_hq_params": (None, [vp, f32, u32, u32, vp, vp, vp, vp, f64, vp]),
    "lw_world_height": (ctypes.c_int32, [vp, ctypes.c_int32, ctypes.c_int32]),
    "lw_world_build": (i32, [vp, ctypes.POINTER(vp)]),
    "lw_world_free": (None, [vp]),
    "lw_dusk_sky_maps": (ctypes.c_long, [vp, u32, u32, u32, u32, vp, sz]),
    "lw_hable_curve": (None, [u32, vp]),
    "lw_encode_thresholds": (None, [vp]),
    "lw_write_pfm": (i32, [ctypes.c_char_p, vp, u32, u32]),
}


def lib():
    """The loaded library; raises when it has not been built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m lumenwright_amd.build`")
        L = ctypes.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib
