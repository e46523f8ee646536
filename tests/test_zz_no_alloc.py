"""lw_hip_render_hq_scene_noalloc, the entry point that allocates nothing, against
lw_hip_render_hq_scene through the C ABI: the images are equal bit for bit on two calls in a
row, the second smaller than the first, and at 720p, where the scratch memory holds fewer
samples per launch than the grow-only buffers do; a frame larger than the scratch memory is
refused with LW_EINVAL before anything runs."""

import numpy as np
import pytest

from lumenwright_amd import frames

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def world():
    cfg = frames.config(1, world_chunks_x=6, world_chunks_z=6, width=160, height=90, spp=8)
    return cfg, frames.make_frame(cfg)


def test_noalloc_render_equals_render_hq_scene(hip_backend, world):
    cfg, fr = world
    small = frames.make_frame(frames.replace(cfg, width=64, height=36, yaw=40.0), fr.scene)
    sc = hip_backend.upload(fr)
    try:
        for params, spp in [(fr.params, 8), (small.params, 3)]:
            want = hip_backend.render_hq_scene(sc, params, spp)
            got = hip_backend.render_hq_scene_noalloc(sc, params, spp)
            print(f"{int(params['frame'][0])}x{int(params['frame'][1])} at {spp} spp: "
                  f"{int(np.all(got.view('<u4') == want.view('<u4'), axis=1).sum())}/{len(got)} "
                  "pixels equal")
            assert np.array_equal(got.view("<u4"), want.view("<u4"))
            assert want.mean() > 0
    finally:
        sc.free()


def test_noalloc_render_at_720p(hip_backend, world):
    cfg, fr = world
    big = frames.make_frame(frames.replace(cfg, width=1280, height=720), fr.scene)
    sc = hip_backend.upload(fr)
    try:
        want = hip_backend.render_hq_scene(sc, big.params, 12)
        got = hip_backend.render_hq_scene_noalloc(sc, big.params, 12)
        t = hip_backend.last_timing()
    finally:
        sc.free()
    assert np.array_equal(got.view("<u4"), want.view("<u4"))
    assert t["samples"] == 1280 * 720 * 12 and t["launches"] >= 2


def test_noalloc_refuses_a_frame_beyond_its_scratch(hip_backend, world):
    import lumenwright_amd as lw
    cfg, fr = world
    p = fr.params.copy()
    p["frame"][0], p["frame"][1] = 4096, 1024  # 4.2 M pixels: 151 MB of the 128 MiB
    sc = hip_backend.upload(fr)
    try:
        with pytest.raises(lw.BackendError) as e:
            hip_backend.render_hq_scene_noalloc(sc, p, 1)
        assert e.value.code == lw.backend.LW_EINVAL and "scratch" in str(e.value)
        # the entry point still works after a refusal
        want = hip_backend.render_hq_scene(sc, fr.params, 2)
        got = hip_backend.render_hq_scene_noalloc(sc, fr.params, 2)
        assert np.array_equal(got.view("<u4"), want.view("<u4"))
    finally:
        sc.free()
