"""Device parity on the feature world (tests/features.py): see-through quads drawn before the
first bounce (integrator.rs:576-591, trace.rs:113-131), paths passing through translucent
surfaces (integrator.rs:645-651), translucent texels of alpha 0 (trace.rs:98), and shadow rays
leaving water or filtered by glass under a lit sun (integrator.rs:319-346).

Bars: trace bit-exact (hit, quad, texel, t bits, cell); render_sample bit-exact in every kernel
form; the render_hq image, the batched images and the differentiable forward at identity
equal the CPU backend's bit for bit; the device's trace count equals the oracle's."""

import os

import numpy as np
import pytest

import features
import rays
from lumenwright_amd import diff, frames
from lumenwright_amd.abi import NO_GROUP, SCENE_GRID

pytestmark = pytest.mark.gpu

MODES = ["persistent", "wavefront", "direct"]


@pytest.fixture(scope="module")
def ffr(oracle):
    return features.feature_frame(oracle)


@pytest.fixture(scope="module")
def form_backends():
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


def test_trace_is_bit_exact_on_the_feature_world(hip_backend, oracle, ffr):
    from test_gpu_parity import trace_fields_equal
    v = ffr.view()
    o, d, tmin, tmax = rays.ray_mix(ffr.params, ffr.scene.quads,
                                    lambda o, d, a, b: oracle.trace(v, o, d, a, b), 20000, seed=31)
    want = oracle.trace(v, o, d, tmin, tmax)
    got = hip_backend.trace(v, o, d, tmin, tmax)
    ok = trace_fields_equal(got, want)
    n_ent = int(np.count_nonzero(want["quad"][want["hit"] == 1] >= len(ffr.scene.volume.quads)))
    print(f"features: {len(o)} rays, {int(want['hit'].sum())} hits ({n_ent} on entity quads), "
          f"{int((~ok).sum())} differ")
    assert ok.all()


@pytest.mark.parametrize("mode", MODES)
def test_render_sample_is_bit_exact_in_every_kernel_form(form_backends, oracle, ffr, mode):
    v = ffr.view()
    rng = np.random.default_rng(41)
    n = 30000
    px = rng.integers(0, 96, n)
    py = rng.integers(0, 54, n)
    s = rng.integers(0, 16, n)
    want = oracle.render_samples(v, px, py, s)
    got = form_backends[mode].render_samples(v, px, py, s)
    same = np.all(got.view("<u4") == want.view("<u4"), axis=1)
    print(f"{mode}: {int(same.sum())}/{n} samples bit-exact")
    assert same.all()


def test_render_hq_equals_the_cpu_backend(hip_backend, oracle, ffr):
    want = oracle.render_hq(ffr.view(), ffr.spp).reshape(-1, 3)
    got = hip_backend.render_hq(ffr)
    same = np.all(got.view("<u4") == want.view("<u4"), axis=1)
    print(f"features render_hq: {int(same.sum())}/{len(same)} pixels bit-identical")
    assert same.all()
    sc = hip_backend.upload(ffr)
    try:
        assert hip_backend.count_traces(sc, ffr.params, 4) == \
            oracle.render_hq_counted(ffr.view(), 4)[1].traces
    finally:
        sc.free()


def test_the_device_built_feature_world_renders_the_same(hip_backend, oracle, ffr):
    # The layout made on the device from the feature volume (overlay, see-through range,
    # glass that occludes nothing) is build_hq's, and renders the same image.
    vol = ffr.scene.volume
    sc, layout, _ = hip_backend.build_scene(vol, ffr.sky)
    try:
        assert np.array_equal(hip_backend.scene_read(sc, SCENE_GRID), ffr.scene.grid)
        assert layout.see_through == tuple(ffr.scene.see_through)
        got = hip_backend.render_hq_scene(sc, ffr.params, 4)
    finally:
        sc.free()
    want = oracle.render_hq(ffr.view(), 4).reshape(-1, 3)
    assert np.array_equal(got.view("<u4"), want.view("<u4"))


def test_batched_feature_views_equal_the_cpu_backend(hip_backend, oracle, ffr):
    base, views = frames.batch_views(features.CONFIG, ffr.scene, n_views=4, seed=8,
                                     sizes=((64, 36), (48, 32)), spps=(4, 8), max_bounces=4)
    sc = hip_backend.upload(base)
    try:
        outs = hip_backend.render_hq_batch(sc, views)
    finally:
        sc.free()
    for (p, spp), img in zip(views, outs):
        want = oracle.render_hq(frames.HqFrame(base.scene, p, spp, base.sky).view(), spp).reshape(-1, 3)
        assert np.array_equal(img.view("<u4"), want.view("<u4"))


def test_differentiable_forward_at_identity_is_render_hq(hip_backend, oracle, ffr):
    s = ffr.scene
    qt = s.quad_type
    p = diff.identity(diff.block_type_groups(s), diff.emitter_groups(s, 8),
                      int(qt[qt != NO_GROUP].max()) + 1, 8)
    sc = hip_backend.upload(ffr)
    try:
        a = hip_backend.render_diff(sc, ffr.params, 4, p.desc())
    finally:
        sc.free()
    want = oracle.render_hq(ffr.view(), 4).reshape(-1, 3)
    assert np.array_equal(a.view("<u4"), want.view("<u4"))
    assert np.array_equal(oracle.render_diff(ffr.view(), 4, p.desc()).reshape(-1, 3).view("<u4"),
                          want.view("<u4"))
