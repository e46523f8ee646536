"""Batched rendering (lw_hip_render_hq_batch) against per-view render_hq on the device and
against the oracle's CPU backend: each view's image must equal the single-view render bit
for bit (same samples, same f64 summation order), in the one persistent launch over
(view, tile) items, in passes with and without hipGraph replay, and in every path-kernel
form."""

import os

import numpy as np
import pytest

from lumenwright_amd import frames

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def batch_setup():
    cfg = frames.config(1, world_chunks_x=6, world_chunks_z=6)
    scene = frames.build_world(cfg.world)
    base, views = frames.batch_views(cfg, scene, n_views=7, seed=5,
                                     sizes=((64, 36), (96, 54), (40, 40), (130, 70)),
                                     spps=(1, 3, 8, 16), max_bounces=8)
    return base, views


def open_backend(env):
    import lumenwright_amd as lw
    saved = {k: os.environ.get(k) for k in env}
    os.environ.update(env)
    try:
        return lw.HipBackend.open(0)
    finally:
        for k, v in saved.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


@pytest.mark.parametrize("env", [{}, {"LW_BATCH_FORM": "passes"},
                                 {"LW_BATCH_FORM": "passes", "LW_BATCH_GRAPHS": "0"},
                                 {"LW_BATCH_FORM": "passes", "LW_BATCH_PASSES": "5"},
                                 {"LW_KERNEL": "persistent"},
                                 {"LW_KERNEL": "persistent", "LW_BATCH_PASSES": "3"},
                                 {"LW_KERNEL": "wavefront"}],
                         ids=["tiles", "graphs", "no-graphs", "5-passes", "persistent",
                              "persistent-3-passes", "wavefront"])
def test_batch_equals_single_view_renders(batch_setup, env):
    base, views = batch_setup
    b = open_backend(env)
    try:
        scene = b.upload(base)
        outs = b.render_hq_batch(scene, views)
        t = b.last_timing()
        assert t["samples"] == sum(int(p["frame"][0]) * int(p["frame"][1]) * s for p, s in views)
        for (p, spp), img in zip(views, outs):
            single = b.render_hq_scene(scene, p, spp)
            assert np.array_equal(img.view("<u4"), single.view("<u4"))
            assert np.isfinite(img).all()
        scene.free()
    finally:
        b.close()


def test_batch_view_matches_the_cpu_backend(batch_setup, oracle):
    base, views = batch_setup
    b = open_backend({})
    try:
        scene = b.upload(base)
        outs = b.render_hq_batch(scene, views[:3])
        for (p, spp), img in zip(views[:3], outs):
            fr = frames.HqFrame(base.scene, p, spp, base.sky)
            want = oracle.render_hq(fr.view(), spp).reshape(-1, 3)
            assert np.array_equal(img.view("<u4"), want.view("<u4"))
        scene.free()
    finally:
        b.close()


def test_batch_rejects_views_that_differ_beyond_the_camera(batch_setup):
    import lumenwright_amd as lw
    base, views = batch_setup
    b = open_backend({})
    try:
        scene = b.upload(base)
        p = views[1][0].copy()
        p["frame"][3] = 2  # max bounces differs
        with pytest.raises(lw.BackendError) as e:
            b.render_hq_batch(scene, [views[0], (p, 4)])
        assert e.value.code == -1
        scene.free()
    finally:
        b.close()


@pytest.mark.parametrize("env", [{"LW_XCD_REMAP": "1"}, {"LW_WAVES": "6"}, {"LW_WAVES": "8"},
                                 {"LW_SPL": "3"}, {"LW_SPL": "8", "LW_WAVES": "6"},
                                 {"LW_SPL": "1"}],
                         ids=["xcd-remap", "6-waves", "8-waves", "3-per-lane", "8-per-lane-6-waves",
                              "1-per-lane"])
def test_launch_variants_give_the_same_image(batch_setup, env):
    # The XCD tile renumbering, the register budgets and the samples per lane change only
    # where and how a sample runs, never its value.
    base, views = batch_setup
    p, spp = views[3]
    ref = open_backend({})
    alt = open_backend(env)
    try:
        s1, s2 = ref.upload(base), alt.upload(base)
        a = ref.render_hq_scene(s1, p, spp)
        b = alt.render_hq_scene(s2, p, spp)
        assert np.array_equal(a.view("<u4"), b.view("<u4"))
        outs = alt.render_hq_batch(s2, views[:4])
        for (pv, sv), img in zip(views[:4], outs):
            assert np.array_equal(img.view("<u4"), ref.render_hq_scene(s1, pv, sv).view("<u4"))
        s1.free()
        s2.free()
    finally:
        ref.close()
        alt.close()


def test_autotune_picks_the_tile_kernels_budget_on_a_large_frame():
    # A 1280 x 720 frame of the config-1 world uses the persistent tile kernel; the tuner
    # times it at 4, 5 and 6 waves per SIMD and keeps the fastest, and the image is unchanged.
    import lumenwright_amd as lw
    cfg = frames.config(1, spp=16)
    frame = frames.make_frame(cfg)
    b = lw.HipBackend.open(0)
    try:
        sc = b.upload(frame)
        before = b.render_hq_scene(sc, frame.params, 16)
        r = b.autotune(sc, frame.params, probe_spp=1)
        tiles = [c for c in r["candidates"] if c["form"] == "tiles"]
        print("tiles:", [(c["waves"], round(c["ms"], 2)) for c in tiles], "best", r.get("best_tiles"))
        assert [c["waves"] for c in tiles] == [4, 5, 6] and all(c["ms"] > 0 for c in tiles)
        assert r["best_tiles"]["ms"] == min(c["ms"] for c in tiles)
        after = b.render_hq_scene(sc, frame.params, 16)
        assert np.array_equal(before.view("<u4"), after.view("<u4"))
        sc.free()
    finally:
        b.close()


def test_autotune_picks_the_fastest_shape_and_keeps_the_images(batch_setup):
    base, views = batch_setup
    p, spp = views[3]
    b = open_backend({})
    try:
        sc = b.upload(base)
        before = b.render_hq_scene(sc, p, spp)
        r = b.autotune(sc, p, probe_spp=2)
        launch = [c["ms"] for c in r["candidates"] if c["form"] != "tiles"]
        assert len(launch) >= 8 and all(m > 0 for m in launch)
        assert r["best"]["ms"] == min(launch)
        tiles = [c["ms"] for c in r["candidates"] if c["form"] == "tiles"]
        assert ("best_tiles" in r) == bool(tiles)
        if tiles:
            assert r["best_tiles"]["ms"] == min(tiles)
        after = b.render_hq_scene(sc, p, spp)
        assert np.array_equal(before.view("<u4"), after.view("<u4"))
        outs = b.render_hq_batch(sc, views[:3])
        for (pv, sv), img in zip(views[:3], outs):
            assert np.array_equal(img.view("<u4"), b.render_hq_scene(sc, pv, sv).view("<u4"))
        sc.free()
    finally:
        b.close()


@pytest.mark.parametrize("spp", [128, 84, 100, 20], ids=["1x1", "2x1", "2x2-7-rounds", "2x2"])
def test_tile_kernel_frame_with_ragged_edges_and_a_partial_sample_group(spp):
    # 641 x 201 pixels. The tile shape follows the spp: at 128 a wave is one pixel's 64 sample
    # slots (2 rounds), at 84 two pixels' 32 (3 rounds, the last with 20 samples), at 100 and
    # 20 four pixels' 16 (the last round with 4 samples); the tiles along the right and bottom
    # edges hold fewer pixels. The persistent tile kernel (one launch) gives the launch ramp's
    # image bit for bit.
    cfg = frames.config(1, width=641, height=201, spp=spp, world_chunks_x=6, world_chunks_z=6)
    frame = frames.make_frame(cfg)
    tiles = open_backend({})
    ramp = open_backend({"LW_FRAME_FORM": "launches"})
    try:
        s1, s2 = tiles.upload(frame), ramp.upload(frame)
        a = tiles.render_hq_scene(s1, frame.params, cfg.spp)
        assert tiles.last_timing()["launches"] == 1
        b = ramp.render_hq_scene(s2, frame.params, cfg.spp)
        assert ramp.last_timing()["launches"] > 1
        assert a.shape == (641 * 201, 3) and np.isfinite(a).all() and (a > 0).any()
        assert np.array_equal(a.view("<u4"), b.view("<u4"))
        s1.free()
        s2.free()
    finally:
        tiles.close()
        ramp.close()


@pytest.mark.parametrize("spp", [64, 96, 80], ids=["one-pixel", "two-pixel", "2x2"])
def test_tile_kernel_frame_matches_the_cpu_backend(oracle, spp):
    # 160 x 128 pixels of the config-1 world with water and emitters: enough tiles and rounds
    # for the persistent tile kernel in each shape (one pixel at 64 spp, two at 96, 2 x 2 at
    # 80). Its image is the CPU backend's (render/cpu/mod.rs) bit for bit.
    cfg = frames.config(1, width=160, height=128, spp=spp, world_chunks_x=6, world_chunks_z=6)
    frame = frames.make_frame(cfg)
    b = open_backend({})
    try:
        sc = b.upload(frame)
        img = b.render_hq_scene(sc, frame.params, spp)
        assert b.last_timing()["launches"] == 1
        want = oracle.render_hq(frame.view(), spp).reshape(-1, 3)
        assert np.array_equal(img.view("<u4"), want.view("<u4"))
        sc.free()
    finally:
        b.close()


def test_a_frame_of_2_24_pixels_takes_two_pixel_tiles():
    # 4096 x 4096 at 64 spp would be 2^24 one-pixel tiles, more than an item can name: the frame
    # takes tiles of two pixels (still one launch) and equals the launch ramp's image.
    cfg = frames.config(1, width=4096, height=4096, spp=64, max_bounces=1, world_chunks_x=6,
                        world_chunks_z=6)
    frame = frames.make_frame(cfg)
    tiles = open_backend({})
    ramp = open_backend({"LW_FRAME_FORM": "launches"})
    try:
        s1, s2 = tiles.upload(frame), ramp.upload(frame)
        a = tiles.render_hq_scene(s1, frame.params, cfg.spp)
        assert tiles.last_timing()["launches"] == 1
        b = ramp.render_hq_scene(s2, frame.params, cfg.spp)
        assert ramp.last_timing()["launches"] > 1
        assert np.array_equal(a.view("<u4"), b.view("<u4"))
        s1.free()
        s2.free()
    finally:
        tiles.close()
        ramp.close()


@pytest.mark.parametrize("env", [{}, {"LW_BATCH_FORM": "passes", "LW_BATCH_PASSES": "2"},
                                 {"LW_BATCH_FORM": "passes", "LW_BATCH_PASSES": "3"}],
                         ids=["tiles", "2-passes", "3-passes"])
def test_batch_passes_with_partial_sample_groups(batch_setup, env):
    # Views of 20, 40, 50, 17, 128 and 84 spp: in passes, a view's samples per pass round up to
    # a multiple of 16 (32 + 8, 32 + 18 for two passes), so passes mix whole 16-sample waves
    # with a remainder; in the one persistent launch the views share one tile shape and an
    # item's last round may be partial. Either way every view equals its single-view render
    # bit for bit.
    base, views = batch_setup
    picked = [(views[0][0], 20), (views[1][0], 40), (views[3][0], 50), (views[2][0], 17),
              (views[4][0], 128), (views[5][0], 84)]
    b = open_backend(env)
    try:
        scene = b.upload(base)
        outs = b.render_hq_batch(scene, picked)
        if env:
            assert b.last_timing()["launches"] == int(env["LW_BATCH_PASSES"])
        for (p, spp), img in zip(picked, outs):
            single = b.render_hq_scene(scene, p, spp)
            assert np.array_equal(img.view("<u4"), single.view("<u4"))
        scene.free()
    finally:
        b.close()


@pytest.mark.parametrize("spps", [(64, 128, 256, 192), (32, 96, 32, 96), (16, 256, 64, 1024)],
                         ids=["one-pixel-tiles", "two-pixel-tiles", "mixed"])
def test_batch_tile_shapes(batch_setup, spps):
    # The batch's one launch takes every view's tile shape when they agree (one pixel's 64
    # samples a round for multiples of 64 spp, two pixels' 32 for 32 and 96) and otherwise the
    # cheapest shape for all; the single-view renders take each view's own. The images agree
    # bit for bit all the same.
    base, views = batch_setup
    picked = [(views[i][0], s) for i, s in zip((3, 1, 0, 2), spps)]
    b = open_backend({})
    try:
        scene = b.upload(base)
        outs = b.render_hq_batch(scene, picked)
        assert b.last_timing()["launches"] == 1
        for (p, spp), img in zip(picked, outs):
            single = b.render_hq_scene(scene, p, spp)
            assert np.array_equal(img.view("<u4"), single.view("<u4"))
            assert np.isfinite(img).all()
        scene.free()
    finally:
        b.close()
