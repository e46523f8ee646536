"""Quick mode on the device (lw_hip_render_quick) against the oracle's restatement of the
reference's CPU rasterizer (crates/lumenwright/src/render/cpu/raster.rs, render/texture.rs),
pixel for pixel: the closed-form scenes of the CPU tests, the synthetic world's frames at the
benchmark sizes, ragged and degenerate frames and meshes, and the call's errors."""

import numpy as np
import pytest

from lumenwright_amd import frames
from lumenwright_amd.abi import QUICK_VERTEX, QuickFrame, QuickMesh

import test_quick as tq

pytestmark = pytest.mark.gpu


def same(hip_backend, oracle, frame, label=""):
    got = hip_backend.render_quick(frame)
    want = oracle.rasterize(frame)
    bad = np.flatnonzero(got != want)
    assert bad.size == 0, (label, bad.size, bad[:4], got[bad[:4]], want[bad[:4]])
    return got


def closed_form_frames():
    f = []
    f.append(("solid", tq.frame_of(tq.mesh_of([tq.facing(-5.0, rect="solid", colour=(255, 128, 64, 255))]))))
    front = tq.facing(-5.0, rect="solid")
    f.append(("back face", tq.frame_of(tq.mesh_of([front[[0, 3, 2, 1]]]))))
    near = tq.facing(-4.0, x0=-10, x1=1, rect="solid")
    far = tq.facing(-6.0, x0=-1, x1=10, rect="green")
    f.append(("depth near first", tq.frame_of(tq.mesh_of([near, far]))))
    f.append(("depth far first", tq.frame_of(tq.mesh_of([far, near]))))
    f.append(("alpha cut", tq.frame_of(tq.mesh_of([tq.facing(-5.0, rect="cut")]))))
    wall = tq.facing(-8.0, rect="green")
    layers = [tq.facing(-9.0, rect="red"), tq.facing(-6.0, rect="red"),
              tq.facing(-5.0, rect="red", colour=(255, 255, 255, 64))]
    f.append(("translucent", tq.frame_of(tq.mesh_of([wall], translucent=layers))))
    f.append(("see-through", tq.frame_of(tq.mesh_of([tq.facing(-3.0, rect="green")],
                                                    see_through=[tq.facing(-9.0, rect="red")]))))
    floor = tq.quad((-30.0, -1.0, 5.0), (60.0, 0, 0), (0, 0, -55.0), rect="solid")
    f.append(("near clip", tq.frame_of(tq.mesh_of([floor]))))
    for which in ("render", "environment"):
        sky = tq.plain_sky(64, 48, fog_colour=(0.2, 0.9, 0.4),
                           render_fog=(4.0, 14.0) if which == "render" else tq.NO_FOG,
                           env_fog=(4.0, 14.0) if which == "environment" else tq.NO_FOG)
        f.append((which + " fog", tq.frame_of(tq.mesh_of([tq.facing(-6.0, rect="solid")]), sky=sky)))
    greys = (40, 90, 140, 190, 240)
    level0, mips = tq.levels_atlas(256, 256, 4, greys)
    for level in (0, 1, 2):
        pmax = 3.0 * 2 ** level
        U, V = pmax * 32 / 256, pmax / 2.5 * 32 / 256
        v = tq.quad((-5.0, -5.0, -10.0), (10.0, 0, 0), (0, 10.0, 0), rect="solid")
        v["uv"] = [[0, 0.05 + V], [U, 0.05 + V], [U, 0.05], [0, 0.05]]
        f.append((f"mip level {level}", tq.frame_of(tq.mesh_of([v]), atlas=level0, atlas_size=(256, 256),
                                                     mips=mips, levels=4, mip_rows=256)))
    return f


def test_closed_form_scenes_are_the_oracles(hip_backend, oracle):
    for label, frame in closed_form_frames():
        same(hip_backend, oracle, frame, label)


def test_random_triangles_are_the_oracles(hip_backend, oracle):
    # Triangles in front of, across and behind the near plane, both windings, all three
    # passes, on a ragged 75 x 41 frame and on a mipmapped atlas with random texels.
    rng = np.random.default_rng(21)
    w, h = 75, 41
    aw, ah = 64, 64
    level0 = rng.integers(0, 2 ** 32, aw * ah, dtype=np.uint64).astype("<u4")
    mips = np.concatenate([rng.integers(0, 2 ** 32, (aw >> l) * (ah >> l), dtype=np.uint64).astype("<u4")
                           for l in (1, 2, 3)])
    n = 600
    v = np.zeros(3 * n, QUICK_VERTEX)
    z = -rng.uniform(-2.0, 30.0, 3 * n)
    v["pos"] = np.stack([rng.uniform(-1.5, 1.5, 3 * n) * np.abs(z) + rng.normal(0, 1, 3 * n),
                         rng.uniform(-1.5, 1.5, 3 * n) * np.abs(z), z], -1).astype("<f4")
    v["uv"] = rng.uniform(-0.1, 1.1, (3 * n, 2)).astype("<f4")
    v["color"] = rng.integers(0, 256, (3 * n, 4))
    idx = np.arange(3 * n, dtype="<u4")
    m = QuickMesh(vertices=v, opaque=idx[:3 * 400], translucent=idx[3 * 400:3 * 550], see_through=idx[3 * 550:])
    sky = tq.plain_sky(w, h, render_fog=(10.0, 25.0))
    f = QuickFrame(mesh=m, atlas_pixels=level0, atlas_size=(aw, ah), atlas_mips=mips, mip_levels=3, mip_rows=40,
                   stars=frames.star_table(), width=w, height=h,
                   view_proj=frames.quick_view_proj(180.0, 0.0, 90.0, w / h, 0.05, 1000.0), sky=sky)
    got = same(hip_backend, oracle, f, "random")
    assert len(np.unique(got)) > 300


@pytest.mark.parametrize("which,size", [(0, None), (0, (75, 41)), (1, None)])
def test_world_frames_are_the_oracles(hip_backend, oracle, which, size):
    cfg = frames.config(which)
    w, h = size or (cfg.width, cfg.height)
    f, _ = frames.quick_frame(cfg, width=w, height=h)
    got = same(hip_backend, oracle, f, f"cfg{which} {w}x{h}")
    t = hip_backend.last_timing()
    print(f"cfg{which} {w}x{h}: {f.mesh.triangle_count()} triangles, {t['trace_ms']:.3f} ms on the device")
    assert t["samples"] == w * h and t["trace_ms"] > 0
    assert len(np.unique(got)) > 100


def test_a_night_frame_is_the_oracles(hip_backend, oracle):
    # The moon's sprite (flipped axes) near the zenith and bright stars, the camera looking up
    # past the config-0 world's hills.
    cfg = frames.config(0, pitch=-75.0)
    f, atlas = frames.quick_frame(cfg, width=320, height=180)
    moon = frames.celestial_direction(10.0)
    d = frames.sky_description(
        sky=(0.02, 0.03, 0.08), fog=(0.04, 0.05, 0.09), sky_fog_end=128.0, render_fog=(115.2, 128.0),
        env_fog=(0.0, 1024.0), sky_disc=True, sunrise=(0.0, 0.0, 0.0, 0.0), sun_angle=190.0,
        moon=(moon, atlas.sun), star_angle=190.0, star_brightness=0.9)
    f.sky = frames.quick_sky_new(frames.camera_rays(cfg.yaw, cfg.pitch, cfg.fov_y, 320, 180), 320, 180, d)
    got = same(hip_backend, oracle, f, "night")
    r, b = got & 0xFF, (got >> 16) & 0xFF
    assert np.count_nonzero((r > 150) & (b < 150)) > 50  # the moon's yellow disc


@pytest.mark.parametrize("w,h", [(1, 1), (1000, 3), (3, 700), (33, 31)])
def test_odd_frame_sizes(hip_backend, oracle, w, h):
    cfg = frames.config(0)
    f, _ = frames.quick_frame(cfg, width=w, height=h)
    same(hip_backend, oracle, f, f"{w}x{h}")


def test_degenerate_meshes(hip_backend, oracle):
    base = tq.frame_of(tq.mesh_of())
    same(hip_backend, oracle, base, "empty")
    q = tq.facing(-5.0, rect="solid")
    # a trailing partial triangle is dropped (raster.rs:199)
    m = QuickMesh(vertices=q, opaque=np.array([0, 1, 2, 0, 2], "<u4"), translucent=np.array([3], "<u4"))
    got = same(hip_backend, oracle, tq.frame_of(m), "partial")
    assert np.any(got != hip_backend.render_quick(base))
    # non-finite and degenerate vertices, a triangle behind the camera, zero-area ones
    v = np.zeros(12, QUICK_VERTEX)
    v["pos"] = [[0, 0, -5], [1, 0, -5], [np.nan, 1, -5],
                [0, 0, -5], [np.inf, 0, -5], [0, 1, -5],
                [0, 0, 5], [1, 0, 5], [0, 1, 5],
                [0, 0, -5], [1, 1, -5], [2, 2, -5]]
    v["uv"] = [[np.nan, 0.1]] * 3 + [[0.1, 0.1]] * 9
    v["color"] = (255, 255, 255, 255)
    m = QuickMesh(vertices=v, opaque=np.arange(12, dtype="<u4"), translucent=np.arange(12, dtype="<u4"))
    same(hip_backend, oracle, tq.frame_of(m), "non-finite")
    # one huge triangle over every tile
    v = np.zeros(3, QUICK_VERTEX)
    v["pos"] = [[-1e4, -1e4, -5], [1e4, -1e4, -5], [0, 1e4, -5]]
    v["uv"] = [[0.1, 0.1]] * 3
    v["color"] = (255, 255, 255, 255)
    same(hip_backend, oracle, tq.frame_of(QuickMesh(vertices=v, opaque=np.arange(3, dtype="<u4")), 300, 200),
         "huge")


def test_render_quick_errors(hip_backend):
    import lumenwright_amd as lw
    ok = tq.frame_of(tq.mesh_of([tq.facing(-5.0, rect="solid")]))
    hip_backend.render_quick(ok)
    bad = tq.frame_of(QuickMesh(vertices=tq.facing(-5.0, rect="solid"), opaque=np.array([0, 1, 4], "<u4")))
    with pytest.raises(lw.BackendError) as e:
        hip_backend.render_quick(bad)  # an index past the vertices
    assert e.value.code == lw.backend.LW_EINVAL
    for change in ({"mip_levels": 17}, {"mip_levels": 4}, {"stars": np.zeros(3, "<u4")}, {"width": 0}):
        f = tq.frame_of(tq.mesh_of([tq.facing(-5.0, rect="solid")]))
        for k, val in change.items():
            setattr(f, k, val)
        with pytest.raises(lw.BackendError) as e:
            hip_backend.render_quick(f)  # 4 levels of an 8 x 8 atlas with no mips given
        assert e.value.code == lw.backend.LW_EINVAL, change
    got = hip_backend.render_quick(ok)  # the handle still works
    assert got.size == 64 * 48
