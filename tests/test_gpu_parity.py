"""Parity of the HIP path with the oracle, through the C ABI, on an MI355X.

Bars (the kernel crate's own outputs, restated by the oracle):
  * trace: bit-exact -- hit or miss, quad, texel, t bits and the final DDA cell;
  * render_sample: bit-exact in every kernel form (the device's e^x follows libm's expf
    method; the two f32 inputs of 2^32 where they part, one ulp apart, are not met here);
  * render_hq: the image equals the CPU backend's bit for bit. The relative error and the
    MAE are printed as diagnostics only.
"""

import os
import types

import numpy as np
import pytest

import rays
import scenes
from lumenwright_amd import frames
from lumenwright_amd.abi import (EMITTER_SAMPLING_ALL, EMITTER_SAMPLING_NONE,
                                 EMITTER_SAMPLING_ONE, EMITTER_SAMPLING_ONE_BLOCK)

pytestmark = pytest.mark.gpu

MODES = ["persistent", "wavefront", "direct"]


@pytest.fixture(scope="module")
def mode_backends():
    """One backend per kernel form (LW_KERNEL is read when a backend opens)."""
    import lumenwright_amd as lw
    out = {}
    old = os.environ.get("LW_KERNEL")
    try:
        for m in MODES:
            os.environ["LW_KERNEL"] = m
            out[m] = lw.HipBackend.open(0)
    finally:
        if old is None:
            os.environ.pop("LW_KERNEL", None)
        else:
            os.environ["LW_KERNEL"] = old
    yield out
    for b in out.values():
        b.close()


def trace_fields_equal(a, b):
    """The traversal result: hit, quad, texel, t bits, final cell; s and tt by value."""
    bad = (a["hit"] != b["hit"]) | (np.any(a["cell"] != b["cell"], axis=1))
    hit = (a["hit"] == 1) & (b["hit"] == 1)
    bad |= hit & ((a["quad"] != b["quad"]) | (a["texel"] != b["texel"])
                  | (a["t"].view("<u4") != b["t"].view("<u4")) | (a["s"] != b["s"])
                  | (a["tt"] != b["tt"]))
    return ~bad


@pytest.fixture(scope="module")
def frame0():
    return frames.make_frame(frames.CONFIGS[0])


@pytest.fixture(scope="module")
def frame1():
    return frames.make_frame(frames.config(1, width=256, height=144, spp=8))


def test_probe_and_name(hip_backend):
    import lumenwright_amd as lw
    desc = lw.probe()
    print("device:", desc, "| backend:", hip_backend.name())
    assert "gfx9" in desc and hip_backend.name().endswith("(hip)")


@pytest.mark.parametrize("which", [0, 1])
def test_trace_is_bit_exact_on_the_config_worlds(hip_backend, oracle, frame0, frame1, which):
    import lumenwright_amd as lw
    fr = frame0 if which == 0 else frame1
    v = fr.view()
    sc = hip_backend.upload(fr)
    try:  # the walk's per-brick octant bits, made on the device from the uploaded grid
        assert np.array_equal(hip_backend.scene_read(sc, lw.abi.SCENE_OCTANT_BOXES), oracle.octant_boxes(v))
    finally:
        sc.free()

    def otrace(o, d, tmin, tmax):
        return oracle.trace(v, o, d, tmin, tmax)

    o, d, tmin, tmax = rays.ray_mix(fr.params, fr.scene.quads, otrace, 20000, seed=11 + which)
    want = oracle.trace(v, o, d, tmin, tmax)
    got = hip_backend.trace(v, o, d, tmin, tmax)
    ok = trace_fields_equal(got, want)
    print(f"cfg{which}: {len(o)} rays, {int(want['hit'].sum())} hits, {int((~ok).sum())} differ")
    assert ok.all(), (np.nonzero(~ok)[0][:10], got[~ok][:3], want[~ok][:3])


def test_trace_meets_the_reference_known_answers(hip_backend):
    # tests.rs:128-195, 606-721 through the device's trace.
    v = scenes.test_scene(False).view()
    o = np.array([[0.5, 5.5, 5.5], [-10.0, 5.5, 5.5], [0.5, 9.5, 5.5], [0.5, 5.5, 5.5],
                  [0.5, 5.5, 5.5], [0.5, 5.5, 5.5]], "<f4")
    d = np.array([[1, 0, 0], [1, 0, 0], [1, 0, 0], [0, 1, 0], [-1, 0, 0], [1, 0, 0]], "<f4")
    tmax = np.array([np.inf] * 5 + [10.0], "<f4")
    h = hip_backend.trace(v, o, d, 0.0, tmax)
    assert list(h["hit"]) == [1, 1, 1, 0, 0, 0]
    assert abs(h["t"][0] - 19.5) < 1e-3 and h["quad"][0] == 3
    assert abs(h["t"][1] - 30.0) < 1e-3 and abs(h["t"][2] - 31.5) < 1e-3
    wall = scenes.wall_scene().view()
    o, d = scenes.wall_rays()
    d = np.array([scenes.normalize32(x) for x in d], "<f4")
    h = hip_backend.trace(wall, o, d)
    x = o[:, 0] + d[:, 0] * h["t"]
    assert np.all((h["hit"] == 1) & (x > 249.99) & (x < 250.01))
    ent = scenes.test_scene(True).view()
    h = hip_backend.trace(ent, [[0.5, 5.5, 5.25], [0.5, 5.5, 5.75]], [[1, 0, 0], [1, 0, 0]])
    assert h["quad"][0] == 6 and abs(h["t"][0] - 8.0) < 1e-3 and abs(h["t"][1] - 19.5) < 1e-3


def test_trace_on_the_test_scenes_matches_the_oracle(hip_backend, oracle):
    rng = np.random.default_rng(3)
    for s in [scenes.test_scene(False), scenes.test_scene(True), scenes.wall_scene(),
              scenes.lamp_scene(EMITTER_SAMPLING_ONE, True)]:
        v = s.view()
        dims = v.params["grid_dims"][:3].astype(np.float64) * 16
        n = 5000
        o = (rng.random((n, 3)) * (dims * 1.4) - dims * 0.2).astype("<f4")
        d = rays._unit(rng.normal(size=(n, 3)))
        d[: n // 5, rng.integers(0, 3)] = 0.0
        d = rays._unit(d)
        want = oracle.trace(v, o, d)
        got = hip_backend.trace(v, o, d)
        assert trace_fields_equal(got, want).all()


def test_the_renderers_walk_stops_when_its_octant_is_empty(hip_backend, oracle):
    # lw_hip_trace reports hits from the walk the renderers run, which gives up on a ray once
    # every brick left toward its octant is empty, and the final cell from the whole walk. The
    # tower's three non-empty bricks sit in different rows and corners (one an overlay only);
    # a third of the rays aim at the cutout quad and the lone cube on the two upper rows, a
    # third rise or stay level, a third go anywhere.
    import lumenwright_amd as lw
    rng = np.random.default_rng(17)
    targets = np.array([[8.5, 53.5, 37.5], [36.5, 69.5, 5.5]])
    for filled in (True, False):
        v = scenes.tower_scene(filled).view()
        sc = hip_backend.upload(scenes.tower_scene(filled))
        try:
            assert np.array_equal(hip_backend.scene_read(sc, lw.abi.SCENE_OCTANT_BOXES), oracle.octant_boxes(v))
        finally:
            sc.free()
        n = 30000
        k = n // 3
        o = rng.random((n, 3)) * [60.0, 120.0, 60.0] - [6.0, 12.0, 6.0]
        d = rng.normal(size=(n, 3))
        d[:k] = targets[rng.integers(0, 2, k)] + rng.uniform(-0.6, 0.6, (k, 3)) - o[:k]
        d[k:2 * k, 1] = np.abs(d[k:2 * k, 1])
        d[k:k + k // 4, 1] = 0.0
        o, d = o.astype("<f4"), rays._unit(d).astype("<f4")
        want = oracle.trace(v, o, d)
        got = hip_backend.trace(v, o, d)
        assert trace_fields_equal(got, want).all(), filled
        if filled:
            hit = got["hit"] == 1
            assert np.count_nonzero(hit & (got["quad"] == 6)) > 100  # the overlay's quad
            assert np.count_nonzero(hit & (got["cell"][:, 1] == 69)) > 100  # the lone cube
        else:
            assert not got["hit"].any()


def sparse_volume(rng, dims, density):
    """Voxels sprinkled at random, so that most 4^3 sub-blocks hold a model and most of their
    voxels are empty (the walk crosses those by the model-free cube jump): model 1 a cube of
    FRONT_ONLY opaque axis quads, model 2 the same cube with the right half of each face cut
    out, model 3 two crossed cutout quads (not axis quads); brick 0 left empty and the last
    brick full."""
    from lumenwright_amd.abi import (ALPHA_CUTOUT, ALPHA_OPAQUE, FLAG_FRONT_ONLY, MODEL,
                                     MODEL_OCCLUDES, QUAD, VoxelVolume)
    nb = int(np.prod(dims))
    vox = np.zeros(nb * 4096, dtype="<u2")
    pick = rng.random(nb * 4096) < density
    vox[pick] = rng.choice(np.array([1, 1, 1, 2, 3], "<u2"), size=int(pick.sum()))
    vox[:4096] = 0
    vox[4096 * (nb - 1):] = 1
    quads = (scenes.cube(ALPHA_OPAQUE | FLAG_FRONT_ONLY) + scenes.cube(ALPHA_CUTOUT | FLAG_FRONT_ONLY)
             + [scenes.quad((0, 1, 0), (1, 0, 1), (0, -1, 0), ALPHA_CUTOUT),
                scenes.quad((1, 1, 0), (-1, 0, 1), (0, -1, 0), ALPHA_CUTOUT)])
    models = [scenes.model(0, 0), scenes.model(0, 6), scenes.model(6, 6), scenes.model(12, 2)]
    return VoxelVolume(dims=dims, voxels=vox, models=np.array(models, dtype=MODEL),
                       model_flags=np.array([0, MODEL_OCCLUDES, 0, 0], np.uint8),
                       quads=np.array(quads, dtype=QUAD), atlas=np.array([0xFFFFFFFF, 0], "<u4"),
                       atlas_size=(2, 1))


@pytest.mark.parametrize("density", [0.01, 0.04])
def test_trace_crosses_empty_voxels_of_occupied_sub_blocks_exactly(hip_backend, oracle, density):
    # The walk crosses an empty voxel of an occupied 4^3 sub-block by the largest model-free
    # cube from it toward the ray's octant (clipped to its brick) in one move; the hit, the
    # texel, the t bits and the final cell must be the cell-by-cell walk's. Rays start inside
    # and outside the grid, a fifth of them along a plane of two axes (a zero step), and a
    # third stop at a finite tmax inside the grid.
    from test_scene_build import view_of
    rng = np.random.default_rng(int(density * 1000))
    dims = (3, 2, 3)
    vol = sparse_volume(rng, dims, density)
    sub = vol.voxels.reshape(-1, 4, 4, 4, 4, 4, 4).transpose(0, 1, 3, 5, 2, 4, 6).reshape(-1, 64)[:-64]
    occupied = (sub != 0).any(axis=1)
    # the case under test dominates: most sub-blocks hold a model, most of their voxels are empty
    assert occupied.mean() > 0.4 and (sub[occupied] == 0).mean() > 0.9
    v = view_of(vol, oracle.build_layout(vol))
    size = np.array(dims, np.float64) * 16
    n = 30000
    o = (rng.random((n, 3)) * size * 1.2 - size * 0.1).astype("<f4")
    d = rng.normal(size=(n, 3))
    d[: n // 5, rng.integers(0, 3)] = 0.0
    d = rays._unit(d)
    tmax = np.where(rng.random(n) < 1 / 3, rng.uniform(0.0, 40.0, n), np.inf).astype("<f4")
    want = oracle.trace(v, o, d, 0.0, tmax)
    got = hip_backend.trace(v, o, d, 0.0, tmax)
    ok = trace_fields_equal(got, want)
    print(f"density {density}: {n} rays, {int(want['hit'].sum())} hits, {int((~ok).sum())} differ")
    assert ok.all(), (np.nonzero(~ok)[0][:10], got[~ok][:3], want[~ok][:3])
    assert 0.1 * n < int(want["hit"].sum()) < 0.95 * n


def cell_cubes_brute_force(grid, off_bricks, stored):
    """For each stored brick, voxel and octant k (sign +1 on axis a where bit a of k is set),
    the largest n <= 8 such that the cube of n^3 voxels from the voxel toward k, clipped to the
    brick, holds no voxel with a model, by counting the models in every such cube (3-d prefix
    sums over the brick mirrored so that k points toward +1 on every axis)."""
    words = grid[off_bricks:off_bricks + stored * 2048].reshape(stored, 2048)
    ids = np.stack([words & 0xFFFF, words >> 16], axis=-1).reshape(stored, 16, 16, 16)  # y, z, x
    full = (ids != 0).astype(np.int64)
    out = np.zeros((stored, 16, 16, 16), np.uint64)
    for k in range(8):
        # mirror the axes (x, z, y of (e, y, z, x)) on which octant k points toward -1
        flip = [a for a, bit in ((3, 1), (2, 4), (1, 2)) if not k & bit]
        f = np.flip(full, axis=flip) if flip else full
        ps = np.zeros((stored, 17, 17, 17), np.int64)
        ps[:, 1:, 1:, 1:] = f.cumsum(1).cumsum(2).cumsum(3)
        i = np.arange(16)
        side = np.zeros(f.shape, np.int64)
        for n in range(1, 9):
            hi = np.minimum(i + n, 16)
            y0, z0, x0 = np.ix_(i, i, i)
            y1, z1, x1 = np.ix_(hi, hi, hi)
            cnt = (ps[:, y1, z1, x1] - ps[:, y0, z1, x1] - ps[:, y1, z0, x1] - ps[:, y1, z1, x0]
                   + ps[:, y0, z0, x1] + ps[:, y0, z1, x0] + ps[:, y1, z0, x0] - ps[:, y0, z0, x0])
            side += (cnt == 0)
        side = np.flip(side, axis=flip) if flip else side
        out |= side.astype(np.uint64) << np.uint64(4 * k)
    return out.astype("<u4").reshape(-1)


def test_the_cell_cube_table_is_the_largest_model_free_cube(hip_backend, oracle, frame0):
    # k_cell_ext against the table's definition, on an uploaded config world and on the sparse
    # worlds both uploaded and built on the device.
    from lumenwright_amd.abi import SCENE_CELL_CUBES
    from test_scene_build import view_of
    cases = [(frame0.view(), frame0.scene.grid, None)]
    for density in (0.01, 0.04):
        vol = sparse_volume(np.random.default_rng(7), (3, 2, 3), density)
        built = oracle.build_layout(vol)
        cases.append((view_of(vol, built), built[0], vol))
    for v, grid, vol in cases:
        off = int(v.params["off"][1])
        top = grid[int(v.params["off"][0]):int(v.params["off"][0]) + int(v.params["grid_dims"][3])]
        stored = int(max([e + 1 for e in top if e != 0xFFFFFFFF and not (e & 0x80000000)], default=0))
        want = cell_cubes_brute_force(np.asarray(grid, "<u4"), off, stored)
        made = [hip_backend.upload(types.SimpleNamespace(view=lambda v=v: v))]
        if vol is not None:
            made.append(hip_backend.build_scene(vol)[0])
        for s in made:
            try:
                got = hip_backend.scene_read(s, SCENE_CELL_CUBES)
            finally:
                s.free()
            assert got.size == want.size and np.array_equal(got, want), np.flatnonzero(got != want)[:8]
        print(f"{stored} stored bricks: table equal")


def sample_set(params, n, spp, seed):
    rng = np.random.default_rng(seed)
    w, h = int(params["frame"][0]), int(params["frame"][1])
    return rng.integers(0, w, n), rng.integers(0, h, n), rng.integers(0, spp, n)


def radiance_report(got, want):
    same = np.all(got.view("<u4") == want.view("<u4"), axis=1)
    denom = np.maximum(np.abs(want), 1e-3)
    rel = np.max(np.abs(got - want) / denom, axis=1)
    return same, rel


def test_render_sample_is_bit_exact_on_config_0(hip_backend, oracle, frame0):
    v = frame0.view()
    px, py, s = sample_set(frame0.params, 40000, 16, 5)
    want = oracle.render_samples(v, px, py, s)
    got = hip_backend.render_samples(v, px, py, s)
    same, rel = radiance_report(got, want)
    print(f"cfg0: {int(same.sum())}/{len(same)} samples bit-exact, worst rel {rel.max():.2e}")
    assert same.all()


def test_render_sample_parity_on_config_1(hip_backend, oracle, frame1):
    v = frame1.view()
    px, py, s = sample_set(frame1.params, 40000, 256, 6)
    want = oracle.render_samples(v, px, py, s)
    got = hip_backend.render_samples(v, px, py, s)
    same, rel = radiance_report(got, want)
    print(f"cfg1: {int(same.sum())}/{len(same)} samples bit-exact, worst rel {rel.max():.2e}")
    assert same.all()


def test_render_hq_matches_the_cpu_backend(hip_backend, oracle, frame0, frame1):
    for fr, name in [(frame0, "cfg0"), (frame1, "cfg1")]:
        spp = 4
        want = oracle.render_hq(fr.view(), spp).reshape(-1, 3)
        got = hip_backend.render_hq(frames.HqFrame(fr.scene, fr.params, spp, fr.sky))
        same = np.all(got.view("<u4") == want.view("<u4"), axis=1)
        mae = float(np.abs(got.astype(np.float64) - want).mean())
        print(f"{name}: {int(same.sum())}/{len(same)} pixels bit-exact, MAE {mae:.3e}")
        assert same.all()


def test_device_image_error_against_the_f64_render(hip_backend, oracle, frame0):
    # The image's error against the f64 render (oracle/lw_oracle_f64.cpp) at config 0 in full.
    want = oracle.render_hq_f64(frame0.view(), frame0.spp).reshape(-1, 3)
    got = hip_backend.render_hq(frame0).astype(np.float64)
    d = np.abs(got - want)
    print(f"cfg0 device image against f64: MAE {d.mean():.3e}, relative {d.mean() / want.mean():.2e}")
    assert d.mean() < 1e-3


def test_device_trace_count_is_the_oracles(hip_backend, oracle, frame0, frame1):
    # lw_hip_count_traces against the instrumented oracle's trace calls for the same render.
    for fr, spp in [(frame0, 4), (frame1, 2)]:
        sc = hip_backend.upload(fr)
        try:
            got = hip_backend.count_traces(sc, fr.params, spp)
        finally:
            sc.free()
        _, c = oracle.render_hq_counted(fr.view(), spp)
        w, h = int(fr.params["frame"][0]), int(fr.params["frame"][1])
        print(f"{w}x{h}x{spp}: {got} traces, {got / (w * h * spp):.3f} per sample")
        assert got == c.traces


def test_water_plane_on_the_device(hip_backend, oracle):
    # Chunky's water plane (integrator.rs:506-534, drawn at no configuration): the device's
    # samples equal the oracle's bit for bit at three incidences, and on the config-0 world
    # with the plane drawn below its terrain and the camera looking out past the grid's edge.
    from test_water_shadow_sun import water_plane_scene
    for cos_i in (0.9, 0.4, 0.15):
        v = water_plane_scene(cos_i).view()
        n = 2048
        want = oracle.render_samples(v, np.zeros(n), np.zeros(n), np.arange(n))
        got = hip_backend.render_samples(v, np.zeros(n), np.zeros(n), np.arange(n))
        assert np.array_equal(got.view("<u4"), want.view("<u4")), cos_i
    fr = frames.make_frame(frames.config(0, yaw=90.0, pitch=10.0, width=160, height=90))
    fr.params["water_plane"] = [62.0, 1.0, 0.0, 0.0]
    v = fr.view()
    px, py, s = sample_set(fr.params, 20000, 16, 8)
    want = oracle.render_samples(v, px, py, s)
    got = hip_backend.render_samples(v, px, py, s)
    same, rel = radiance_report(got, want)
    assert same.all(), (int((~same).sum()), rel.max())


def test_render_hq_kat_scenes_on_the_device(hip_backend):
    # tests.rs:475-492: the lamp seen from the camera is 0.5 * 13 in every sample.
    s = scenes.lamp_scene(EMITTER_SAMPLING_ONE, False)
    p = s.params
    p["frame"] = [1, 1, 0, 1]
    p["cam_pos"] = [7.5, 2.0, 7.5, 0.01]
    p["cam_fwd"] = [0.0, 1.0, 0.0, 1.0]
    p["cam_right"] = [1.0, 0.0, 0.0, 0.0]
    p["cam_up"] = [0.0, 0.0, -1.0, 0.0]
    c = hip_backend.render_samples(s.view(), np.zeros(16), np.zeros(16), np.arange(16))
    assert np.all(np.abs(c[:, :2] - 6.5) < 1e-4)
    # tests.rs:804-831: lens and light maps.
    s = scenes.test_scene(False)
    p = s.params
    p["sky_map"] = [4.0, 2.0, 2.0, 2.0]
    s.sky = np.concatenate([np.full(24, 5.0), np.full(12, 0.3)]).astype("<f4")
    p["frame"] = [1, 1, 0, 1]
    p["cam_pos"] = [40.25, 20.0, 8.25, 1e-3]
    p["cam_right"] = [1.0, 0.0, 0.0, 0.0]
    p["sun_dir"] = [0.0, 1.0, 0.0, 0.9995]
    p["cam_fwd"] = [0.0, -1.0, 0.0, 1.0]
    p["cam_up"] = [0.0, 0.0, -1.0, 0.0]
    c = hip_backend.render_samples(s.view(), np.zeros(16), np.zeros(16), np.arange(16))
    assert np.all(np.abs(c - 0.3) < 1e-4)


def test_render_hq_is_deterministic_and_splits_exactly(hip_backend, frame1):
    fr = frames.HqFrame(frame1.scene, frame1.params, 8, frame1.sky)
    a = hip_backend.render_hq(fr)
    b = hip_backend.render_hq(fr)
    assert np.array_equal(a.view("<u4"), b.view("<u4"))
    scene = hip_backend.upload(fr)
    c = hip_backend.render_hq_scene(scene, fr.params, 8)
    assert np.array_equal(a.view("<u4"), c.view("<u4"))
    t = hip_backend.last_timing()
    assert t["samples"] == 256 * 144 * 8 and t["launches"] >= 1
    scene.free()


@pytest.mark.parametrize("mode", MODES)
def test_every_kernel_form_matches_the_oracle(mode_backends, oracle, frame0, frame1, mode):
    b = mode_backends[mode]
    for fr, spp, seed in [(frame0, 16, 21), (frame1, 256, 22)]:
        v = fr.view()
        px, py, s = sample_set(fr.params, 20000, spp, seed)
        want = oracle.render_samples(v, px, py, s)
        got = b.render_samples(v, px, py, s)
        same, rel = radiance_report(got, want)
        print(f"{mode}: {int(same.sum())}/{len(same)} bit-exact, worst rel {rel.max():.2e}")
        assert same.all()
    img = b.render_hq(frames.HqFrame(frame1.scene, frame1.params, 3, frame1.sky))
    want = oracle.render_hq(frame1.view(), 3).reshape(-1, 3)
    assert np.array_equal(img.view("<u4"), want.view("<u4"))


def lamp_camera(s):
    p = s.params
    p["frame"] = [8, 8, 0, 3]
    p["cam_pos"] = [8.0, 3.2, 8.0, 0.9]
    p["cam_fwd"] = [0.0, -1.0, 0.0, 1.0]
    p["cam_right"] = [1.0, 0.0, 0.0, 0.0]
    p["cam_up"] = [0.0, 0.0, -1.0, 0.0]
    return s


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("strategy", [EMITTER_SAMPLING_NONE, EMITTER_SAMPLING_ONE,
                                      EMITTER_SAMPLING_ONE_BLOCK, EMITTER_SAMPLING_ALL])
@pytest.mark.parametrize("decoy", [False, True])
def test_emitter_sampling_strategies_are_bit_exact(mode_backends, oracle, mode, strategy, decoy):
    # The lamp scenes of tests.rs:310-395, seen from above the floor: emitter sampling of
    # each strategy (integrator.rs:418-469), through every kernel form.
    s = lamp_camera(scenes.lamp_scene(strategy, decoy))
    v = s.view()
    n = 8 * 8 * 64
    px = np.tile(np.repeat(np.arange(8), 8), 64)
    py = np.tile(np.arange(8), 8 * 64)
    smp = np.repeat(np.arange(64), 64)
    want = oracle.render_samples(v, px, py, smp)
    got = mode_backends[mode].render_samples(v, px, py, smp)
    assert np.array_equal(got.view("<u4"), want.view("<u4")), (strategy, decoy)
    assert n == len(got) and want.mean() > 0.0


def test_the_persistent_frame_form_equals_the_launch_ramp(hip_backend, oracle):
    # A frame with 16 spp or more and enough 4 x 4 tiles to give every resident wave several
    # (the 720p frame) renders in one persistent launch, every tile's samples summed in sample
    # order in one wave; the launch ramp with staging and an f64 accumulation pass
    # (LW_FRAME_FORM=launches) and the CPU backend give the same image bit for bit, and
    # progress climbs to (spp, spp).
    import lumenwright_amd as lw
    fr = frames.make_frame(frames.config(1, spp=20))
    calls = []
    got = hip_backend.render_hq(fr, progress=lambda d, t: calls.append((d, t)))
    t = hip_backend.last_timing()
    assert t["launches"] == 1 and t["samples"] == 1280 * 720 * 20
    assert calls and calls[-1] == (20, 20) and all(t == 20 for _, t in calls)
    assert all(a[0] <= b[0] for a, b in zip(calls, calls[1:]))
    old = os.environ.get("LW_FRAME_FORM")
    os.environ["LW_FRAME_FORM"] = "launches"
    try:
        ramp = lw.HipBackend.open(0)
    finally:
        if old is None:
            os.environ.pop("LW_FRAME_FORM", None)
        else:
            os.environ["LW_FRAME_FORM"] = old
    try:
        want = ramp.render_hq(fr)
        assert ramp.last_timing()["launches"] > 1
    finally:
        ramp.close()
    assert np.array_equal(got.view("<u4"), want.view("<u4"))
    cpu = oracle.render_hq(fr.view(), 20).reshape(-1, 3)
    assert np.array_equal(got.view("<u4"), cpu.view("<u4"))
