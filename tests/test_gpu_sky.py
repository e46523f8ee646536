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
