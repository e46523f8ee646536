"""The scene layout made on the device (lw_hip_scene_build) against the oracle's restatement
of build_hq (crates/lumenwright/src/scene/voxel.rs:305-491, emitter_list :542-607): the grid,
the emitter list and the emitter grid equal word for word; and frames rendered from a
device-built scene equal those rendered from the same scene built on the host."""

import dataclasses

import numpy as np
import pytest

from lumenwright_amd import frames
from lumenwright_amd.abi import (MODEL, MODEL_EMITS, MODEL_OCCLUDES, QUAD, SCENE_EMITTER_GRID,
                                 SCENE_EMITTERS, SCENE_GRID, SCENE_OCTANT_BOXES, SCENE_QUADS,
                                 VoxelVolume)
from test_scene_build import (DEEP, emitter_volume, kernel_test_volume, random_entity_quads,
                              view_of)

pytestmark = pytest.mark.gpu


def built_equals_oracle(b, oracle, vol, sky=None):
    sc, layout, stats = b.build_scene(vol, sky)
    try:
        g = b.scene_read(sc, SCENE_GRID)
        em = b.scene_read(sc, SCENE_EMITTERS)
        eg = b.scene_read(sc, SCENE_EMITTER_GRID)
        qd = b.scene_read(sc, SCENE_QUADS)
        boxes = b.scene_read(sc, SCENE_OCTANT_BOXES)
    finally:
        sc.free()
    og, oem, oeg, info = oracle.build_layout(vol)
    # the walk's per-brick octant bits, made on the device from the built grid
    assert np.array_equal(boxes, oracle.octant_boxes(view_of(vol, (og, oem, oeg, info))))
    assert g.size == og.size and np.array_equal(g, og), np.flatnonzero(g[:og.size] != og[:g.size])[:8]
    assert np.array_equal(em.view("u1"), oem.view("u1"))
    assert np.array_equal(eg, oeg)
    quads = np.concatenate([vol.quads, vol.entity_quads]) if len(vol.entity_quads) else vol.quads
    if len(quads) == 0:
        quads = np.zeros(1, dtype=QUAD)
    assert np.array_equal(qd.view("u1"), np.ascontiguousarray(quads, dtype=QUAD).view("u1"))
    assert layout.offsets == info["offsets"] and layout.see_through == info["see_through"]
    assert layout.emitter_count == info["emitter_count"]
    for k in ("bricks", "uniform_bricks", "buried_bricks", "overlay_bricks"):
        assert stats[k] == info[k], k
    assert stats["emitters"] == info["emitter_count"] and stats["grid_len"] == og.size
    return stats


@pytest.mark.parametrize("which", [0, 1])
def test_the_config_worlds_build_as_the_oracle_builds_them(hip_backend, oracle, which):
    cfg = frames.CONFIGS[which]
    scene = frames.build_world(cfg.world, volume=True)
    stats = built_equals_oracle(hip_backend, oracle, scene.volume)
    print(f"config {which}: {stats['bricks']} stored, {stats['uniform_bricks']} uniform, "
          f"{stats['buried_bricks']} buried bricks, {stats['emitters']} emitters; "
          f"{stats['kernel_ms']:.2f} ms on the device, {stats['wall_ms']:.1f} ms in all")


@pytest.mark.parametrize("given_occlusion", [True, False])
def test_buried_and_uniform_bricks(hip_backend, oracle, given_occlusion):
    vol = frames.build_world(DEEP, volume=True).volume
    if not given_occlusion:
        vol = dataclasses.replace(vol, brick_occludes=None)
    stats = built_equals_oracle(hip_backend, oracle, vol)
    assert stats["buried_bricks"] > 0 and stats["uniform_bricks"] > 0


def test_the_overlay_and_its_see_through_range(hip_backend, oracle):
    rng = np.random.default_rng(9)
    dims = (3, 2, 2)
    vol = dataclasses.replace(emitter_volume([(5, 5, 5), (40, 20, 30), (0, 0, 0)], dims),
                              entity_quads=random_entity_quads(rng, 300, np.array(dims) * 16),
                              see_through_count=11)
    stats = built_equals_oracle(hip_backend, oracle, vol)
    assert stats["overlay_bricks"] > 0 and stats["entity_quads"] == 300


def random_volume(rng, dims, n_models, fill):
    nb = int(np.prod(dims))
    vox = np.where(rng.random(nb * 4096) < fill, rng.integers(1, n_models, nb * 4096), 0)
    vox = vox.astype("<u2")
    models = np.zeros(n_models, dtype=MODEL)
    models["quad_offset"] = rng.integers(0, 4, n_models)
    models["quad_count"] = rng.integers(0, 3, n_models)
    flags = rng.integers(0, 4, n_models).astype(np.uint8)
    flags[0] = 0
    return VoxelVolume(dims=dims, voxels=vox, models=models, model_flags=flags,
                       quads=np.zeros(8, dtype=QUAD), atlas=np.array([0xFFFFFFFF], "<u4"),
                       atlas_size=(1, 1))


@pytest.mark.parametrize("case", ["empty", "one-uniform", "all-solid", "sparse", "dense",
                                  "kernel-test", "emitter-grid-test"])
def test_edge_volumes(hip_backend, oracle, case):
    rng = np.random.default_rng(len(case))
    if case == "empty":
        vol = dataclasses.replace(random_volume(rng, (2, 1, 2), 3, 0.0))
    elif case == "one-uniform":
        vol = random_volume(rng, (1, 1, 1), 3, 0.0)
        vol.voxels[:] = 2
    elif case == "all-solid":  # the middle brick of 3x3x3 is buried
        vol = random_volume(rng, (3, 3, 3), 2, 0.0)
        vol.voxels[:] = 1
        vol.model_flags[1] = MODEL_OCCLUDES
    elif case == "sparse":
        vol = random_volume(rng, (5, 3, 4), 700, 0.01)
    elif case == "dense":
        vol = random_volume(rng, (4, 4, 4), 40, 0.9)
    elif case == "kernel-test":
        vol = kernel_test_volume(True)
    else:
        vol = emitter_volume([(25, 3, 0), (4, 40, 17), (26, 3, 1)], (2, 3, 2))
    stats = built_equals_oracle(hip_backend, oracle, vol)
    if case == "all-solid":
        assert stats["buried_bricks"] == 1


def test_a_device_built_config_world_renders_as_the_uploaded_one(hip_backend):
    cfg = frames.config(1, width=160, height=90, spp=4, world_chunks_x=6, world_chunks_z=6)
    scene = frames.build_world(cfg.world, volume=True)
    frame = frames.make_frame(cfg, scene)
    built, layout, _ = hip_backend.build_scene(scene.volume, frame.sky)
    up = hip_backend.upload(frame)
    try:
        assert layout.offsets == tuple(scene.offsets)
        a = hip_backend.render_hq_scene(built, frame.params, cfg.spp)
        c = hip_backend.render_hq_scene(up, frame.params, cfg.spp)
        assert np.array_equal(a.view("<u4"), c.view("<u4"))
        # a second render checks the params against the layout fixed by the first
        a2 = hip_backend.render_hq_scene(built, frame.params, 1)
        assert np.isfinite(a2).all()
    finally:
        built.free()
        up.free()


def test_a_device_built_overlay_renders_as_the_oracle_scene(hip_backend, oracle):
    vol = kernel_test_volume(True)
    built, layout, _ = hip_backend.build_scene(vol)
    try:
        v = view_of(vol, oracle.build_layout(vol))
        p = v.params
        p["frame"] = [24, 16, 0, 3]
        p["cam_pos"] = [0.5, 5.5, 5.5, 0.6]
        p["cam_fwd"] = [1.0, 0.0, 0.0, 1.5]
        p["cam_right"] = [0.0, 0.0, 1.0, 0.0]
        p["cam_up"] = [0.0, 1.0, 0.0, 0.0]
        p["sun_dir"] = [0.3, 0.9, 0.1, 0.9995]
        v = view_of(vol, oracle.build_layout(vol), params=p)
        got = hip_backend.render_hq_scene(built, v.params, 8)
        want = oracle.render_hq(v, 8).reshape(-1, 3)
        assert np.array_equal(got.view("<u4"), want.view("<u4"))
    finally:
        built.free()


def test_bad_volumes_and_mismatched_params_are_errors(hip_backend):
    import lumenwright_amd as lw
    rng = np.random.default_rng(1)
    vol = random_volume(rng, (2, 1, 1), 5, 0.5)
    bad = dataclasses.replace(vol, voxels=vol.voxels.copy())
    bad.voxels[77] = 9  # past models_len
    with pytest.raises(lw.BackendError):
        hip_backend.build_scene(bad)
    air = dataclasses.replace(vol, model_flags=vol.model_flags.copy())
    air.model_flags[0] = MODEL_EMITS
    with pytest.raises(lw.BackendError):
        hip_backend.build_scene(air)
    with pytest.raises(lw.BackendError):
        hip_backend.build_scene(dataclasses.replace(vol, dims=(1 << 12, 1 << 12, 1)))
    sc, layout, _ = hip_backend.build_scene(vol)
    try:
        cfg = frames.config(0, width=8, height=8, spp=1)
        frame = frames.make_frame(cfg)  # another scene's layout
        with pytest.raises(lw.BackendError):
            hip_backend.render_hq_scene(sc, frame.params, 1)
    finally:
        sc.free()
