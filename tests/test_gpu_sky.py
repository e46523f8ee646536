"""Sky preparation on the device against the host code and the oracle: the sky cache
(build_sky_cache, crates/lumenwright-kernel/src/integrator.rs:133-148) and the blend of
Chunky's sky into the game's maps (blend_chunky_sky, crates/lumenwright/src/render/
sky.rs:836-851), both bit for bit."""

import numpy as np
import pytest

from lumenwright_amd import frames
from lumenwright_amd.abi import BoundView

pytestmark = pytest.mark.gpu


def chunky_params(which: int, res: float = 128.0):
    frame = frames.make_frame(frames.config(which, width=8, height=8, spp=1))
    p = frame.params.copy()
    p["misc_f"][2] = res
    p["sky_map"] = [0, 0, 0, 0]
    return frame, p


@pytest.mark.parametrize("which", [0, 3])
@pytest.mark.parametrize("res", [128.0, 37.0, 1.0, 0.0])
def test_the_sky_cache_on_the_device_is_the_hosts_and_the_oracles(hip_backend, oracle, which, res):
    frame, p = chunky_params(which, res)
    got = hip_backend.sky_cache(p)
    host = frames.sky_cache(p)
    assert np.array_equal(got.view("<u4"), host.view("<u4"))
    if res:
        s = frame.scene
        v = BoundView(p, s.grid, s.models, s.quads, s.atlas, s.emitters, s.emitter_grid, None)
        want = oracle.build_sky_cache(v, int(res))
        assert np.array_equal(got.view("<u4"), np.asarray(want, "<f4").reshape(-1).view("<u4"))
    else:
        assert got.size == 0


@pytest.mark.parametrize("weight", [0.0, 0.37, 1.0])
def test_the_blend_on_the_device_is_the_oracles(hip_backend, oracle, weight):
    # The game look's full-size maps (2048 x 1024 lens, 360 x 180 light) and Chunky's sky of
    # config 0 (the day sky) as the blend's source.
    frame, p = chunky_params(0)
    sun = frames.sun_direction_from_angle(frames.celestial_angle(12800))
    maps = frames.dusk_sky_maps(sun)
    sizes = [2048, 1024, 360, 180]
    p["sky_map"] = sizes
    got = hip_backend.blend_chunky_sky(p, weight, maps)
    s = frame.scene
    v = BoundView(p, s.grid, s.models, s.quads, s.atlas, s.emitters, s.emitter_grid, frame.sky)
    want = oracle.blend_chunky_sky(v, weight, sizes, maps)
    bad = np.flatnonzero(got.view("<u4") != want.view("<u4"))
    assert bad.size == 0, (bad.size, bad[:4], got[bad[:4]], want[bad[:4]])
    t = hip_backend.last_timing()
    print(f"blend of {t['samples']} texels: {t['trace_ms']:.3f} ms")


def test_sky_preparation_errors(hip_backend):
    import lumenwright_amd as lw
    _, p = chunky_params(0)
    p["sky_map"] = [4, 2, 2, 2]
    maps = np.zeros((8 + 4) * 3, dtype="<f4")
    with pytest.raises(lw.BackendError):
        hip_backend.blend_chunky_sky(p, 0.5, maps[:-3])  # not the maps' size
    p["misc_f"][2] = 0.0
    with pytest.raises(lw.BackendError):
        hip_backend.blend_chunky_sky(p, 0.5, maps)  # no cache to blend from
    p["misc_f"][2] = -3.0
    with pytest.raises(lw.BackendError):
        hip_backend.sky_cache(p)


def moonlit_description():
    """A night sky: the moon drawn (flipped sprite axes), stars bright, no glow."""
    moon = frames.celestial_direction(200.0)
    return frames.sky_description(
        sky=(0.02, 0.03, 0.08), fog=(0.04, 0.05, 0.09), sky_fog_end=256.0,
        render_fog=(230.4, 256.0), env_fog=(0.0, 1024.0), sky_disc=True,
        sunrise=(0.0, 0.0, 0.0, 0.0), sun_angle=200.0, moon=(moon, (2.0, 0.0, 2.0, 2.0)),
        star_angle=200.0, star_brightness=0.9)


@pytest.mark.parametrize("case", ["dusk", "night"])
def test_the_game_sky_maps_on_the_device_are_the_oracles(hip_backend, oracle, case):
    from test_game_sky import dusk_camera, dusk_description
    cam = dusk_camera()
    # 4 x 3 texels: the sprites' rects in the first two rows, a spare row and column, since a
    # texel coordinate can round up to a rect's far edge
    atlas = np.array([0xFFFFFFFF, 0xFF8040C0, 0x80FFFFFF, 0xFF000000,
                      0xFFFFFFFF, 0xFF8040C0, 0x80FFFFFF, 0xFF000000,
                      0xFF0000FF, 0xFF00FF00, 0xFFFF0000, 0xFFFFFFFF], "<u4")
    if case == "dusk":
        lens_d, light_d, factor = dusk_description(True), dusk_description(False), 0.378182
    else:
        lens_d, factor = moonlit_description(), 0.24
        light_d = moonlit_description()
        light_d.has_moon, light_d.star_brightness = 0, 0.0
    lens = frames.quick_sky_new(cam, 16, 9, lens_d)
    light = frames.quick_sky_new(cam, 16, 9, light_d)
    colour = np.array([0x92, 0x92, 0xFF], "<f4") / np.float32(255.0)
    got, info = hip_backend.game_sky_maps(lens, light, factor, colour, atlas, 4, frames.star_table())
    want, winfo = oracle.game_sky_maps(lens, light, factor, colour, atlas, 4)
    print(f"{case}: {info['kernel_ms']:.3f} ms on the device (table set-up {info['setup_ms']:.0f} ms), "
          f"scale {info['scale']}, veil {info['veil']}")
    for k in ("light", "mean", "veil"):
        assert np.array_equal(np.array(info[k], "<f4"), np.array(winfo[k], "<f4")), k
    assert np.float32(info["scale"]) == np.float32(winfo["scale"])
    bad = np.flatnonzero(got.view("<u4") != want.view("<u4"))
    assert bad.size == 0, (bad.size, bad[:6] // 3, got[bad[:6]], want[bad[:6]])
