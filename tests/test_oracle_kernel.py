"""Pins the oracle (the C++ restatement of crates/lumenwright-kernel) to the reference's own
known-answer tests, crates/lumenwright-kernel/src/tests.rs, restated one for one, and to
brute-force f64 checks the mathematics fixes."""

import numpy as np
import pytest

import scenes
from lumenwright_amd.abi import (ALPHA_CUTOUT, EMITTER_SAMPLING_ALL, EMITTER_SAMPLING_NONE,
                                 EMITTER_SAMPLING_ONE, EMITTER_SAMPLING_ONE_BLOCK,
                                 FLAG_FRONT_ONLY, f32_bits)

INF = np.float32(np.inf)


def one(o, d, view, oracle, tmin=0.0, tmax=np.inf):
    h = oracle.trace(view, np.array([o], "<f4"), np.array([d], "<f4"), tmin, tmax)[0]
    return h if h["hit"] else None


def test_sizes_match_the_reference_layout(oracle):
    # trace.rs:327-333
    L = oracle.lib()
    assert [L.lwo_sizeof(i) for i in range(5)] == [432, 80, 16, 32, 40]


def test_ray_skips_empty_brick_and_hits_cube(oracle):
    # tests.rs:128-136
    v = scenes.test_scene(False).view()
    h = one([0.5, 5.5, 5.5], [1, 0, 0], v, oracle)
    assert h is not None and abs(h["t"] - 19.5) < 1e-3 and h["quad"] == 3
    assert list(h["cell"]) == [20, 5, 5]


def test_ray_from_outside_grid_enters_it(oracle):
    # tests.rs:138-151
    v = scenes.test_scene(False).view()
    h = one([-10.0, 5.5, 5.5], [1, 0, 0], v, oracle)
    assert h is not None and abs(h["t"] - 30.0) < 1e-3


def test_ray_reaches_uniform_brick(oracle):
    # tests.rs:153-160
    v = scenes.test_scene(False).view()
    h = one([0.5, 9.5, 5.5], [1, 0, 0], v, oracle)
    assert h is not None and abs(h["t"] - 31.5) < 1e-3


def test_ray_misses_everything(oracle):
    # tests.rs:162-179
    v = scenes.test_scene(False).view()
    assert one([0.5, 5.5, 5.5], [0, 1, 0], v, oracle) is None
    assert one([0.5, 5.5, 5.5], [-1, 0, 0], v, oracle) is None
    assert one([0.5, 5.5, 5.5], [1, 0, 0], v, oracle, tmax=10.0) is None


def test_oblique_rays_hit_cube_top(oracle):
    # tests.rs:181-195
    v = scenes.test_scene(False).view()
    o = np.array([18.0, 9.0, 3.0], "<f4")
    d = scenes.normalize32(np.array([20.5, 6.0, 5.5], "<f4") - o)
    h = one(o, d, v, oracle)
    p = o + d * h["t"]
    assert abs(p[1] - 6.0) < 1e-3 and h["quad"] == 4


def test_entity_overlay_quads_are_hit_and_alpha_tested(oracle):
    # tests.rs:197-222
    v = scenes.test_scene(True).view()
    h = one([0.5, 5.5, 5.25], [1, 0, 0], v, oracle)
    assert abs(h["t"] - 8.0) < 1e-3 and h["quad"] == 6
    h = one([0.5, 5.5, 5.75], [1, 0, 0], v, oracle)
    assert abs(h["t"] - 19.5) < 1e-3


def test_front_only_quads_are_met_from_the_front_only(oracle):
    # tests.rs:224-240
    s = scenes.test_scene(True)
    s.quads[6]["p0"][3] = f32_bits(ALPHA_CUTOUT | FLAG_FRONT_ONLY)
    v = s.view()
    h = one([0.5, 5.5, 5.25], [1, 0, 0], v, oracle, tmax=16.0)
    assert h is not None and h["quad"] == 6
    assert one([12.0, 5.5, 5.25], [-1, 0, 0], v, oracle) is None


def test_a_see_through_quad_is_found_behind_what_hides_it(oracle):
    # tests.rs:242-259
    s = scenes.test_scene(True)
    s.params["see_through"] = [6, 1, 0, 0]
    v = s.view()
    assert oracle.is_see_through(v, 6) and not oracle.is_see_through(v, 5)
    assert not oracle.is_see_through(v, 7)
    o, d = np.array([25.5, 5.5, 5.25], "<f4"), np.array([-1.0, 0, 0], "<f4")
    h = one(o, d, v, oracle)
    assert h["quad"] == 2
    st = oracle.see_through_hit(v, o, d, h["t"])
    assert st is not None and st["quad"] == 6 and abs(st["t"] - 17.0) < 1e-3
    assert oracle.see_through_hit(v, [25.5, 5.5, 5.75], d, h["t"]) is None
    s.params["see_through"] = [0, 0, 0, 0]
    assert oracle.see_through_hit(s.view(), o, d, h["t"]) is None


def sky_scene():
    """tests.rs:261-284: the empty-scene sky setup."""
    s = scenes.test_scene(False)
    p = s.params
    p["frame"] = [4, 4, 0, 4]
    p["cam_pos"] = [0.5, 12.0, 0.5, 0.5]
    p["cam_fwd"] = [0.0, 1.0, 0.0, 1.0]
    p["cam_right"] = [1.0, 0.0, 0.0, 0.0]
    p["cam_up"] = [0.0, 0.0, -1.0, 0.0]
    p["sky_sun"] = [0.5, np.float32(np.sqrt(0.5)), 0.5, 1.0]
    p["sky_zenith"] = scenes.PREETHAM_45_ZENITH
    for k, name in enumerate(["perez_a", "perez_b", "perez_c", "perez_d", "perez_e"]):
        p[name] = scenes.PREETHAM_PEREZ[k]
    p["sun_dir"] = [0.0, -1.0, 0.0, 1.0]
    p["misc_f"][2] = 128.0
    return s


def test_empty_scene_shows_sky(oracle):
    # tests.rs:261-291
    s = sky_scene()
    s.sky = oracle.build_sky_cache(s.view(), 128)
    v = s.view()
    px, py = np.meshgrid(np.arange(4), np.arange(4))
    c = oracle.render_samples(v, px.ravel(), py.ravel(), np.zeros(16))
    assert np.all(c[:, 2] > c[:, 0]) and np.all(c[:, 2] > 0.3), c


def test_rng_matches_reference_values(oracle):
    # tests.rs:293-298
    assert oracle.pcg_hash(0) == 129708002
    assert oracle.pcg_hash(1) == 2831084092


def pcg_py(v):
    """rng.rs:5-9 in Python integers: the closed form the KAT above pins."""
    state = (v * 747796405 + 2891336453) & 0xFFFFFFFF
    word = (((state >> ((state >> 28) + 4)) ^ state) * 277803737) & 0xFFFFFFFF
    return (word >> 22) ^ word


def test_rng_stream_is_the_pcg_composition(oracle):
    # rng.rs:23-34: state = pcg(pix ^ pcg(sample + pcg(seed))); draw (pcg(state) >> 8) * 2^-24
    rs = np.random.default_rng(5)
    for pix, sample, seed in rs.integers(0, 2**32, size=(200, 3)):
        pix, sample, seed = int(pix), int(sample), int(seed)
        st = pcg_py(pix ^ pcg_py((sample + pcg_py(seed)) & 0xFFFFFFFF))
        assert oracle.rng_new(pix, sample, seed) == st
        draws, _ = oracle.rng_draws(st, 5)
        for k in range(5):
            st = pcg_py(st)
            assert draws[k] == np.float32((st >> 8) / 16777216.0)


def floor_emitter_light(s, draws):
    # tests.rs:400-411
    v = s.view()
    pos = np.array([7.3, 1.0, 7.6], "<f4")
    n = np.array([0.0, 1.0, 0.0], "<f4")
    origin = (pos + n * np.float32(1e-3)).astype("<f4")
    st = oracle_rng_new(3, 0, 11)
    out = ORACLE.sample_emitters(v, pos, origin, n, st, draws)
    return float(np.sum(out[:, 1].astype(np.float64)) / draws)


def floor_emitter_expectation():
    # tests.rs:413-429 (vectorised f64 quadrature)
    px, py, pz = 7.3, 1.001, 7.6
    k = 400
    g = 7.0 + (np.arange(k) + 0.5) / k
    x, z = np.meshgrid(g, g, indexing="ij")
    d = np.stack([x - px, np.full_like(x, 4.0 - py), z - pz])
    d2 = (d * d).sum(0)
    acc = ((d[1] / np.sqrt(d2)) / np.maximum(d2, 1.0) / (k * k)).sum()
    return np.pi * scenes.LAMP_INTENSITY * scenes.LAMP_EMITTANCE * acc


ORACLE = None


def oracle_rng_new(*a):
    return ORACLE.rng_new(*a)


@pytest.fixture(autouse=True)
def _bind(oracle):
    global ORACLE
    ORACLE = oracle


def test_emitter_sampling_is_chunkys_estimate():
    # tests.rs:434-453
    want = floor_emitter_expectation()
    for strategy in [EMITTER_SAMPLING_ONE, EMITTER_SAMPLING_ONE_BLOCK, EMITTER_SAMPLING_ALL]:
        got = floor_emitter_light(scenes.lamp_scene(strategy, False), 200_000)
        assert abs((got - want) / want) < 0.01, (strategy, got, want)
    assert floor_emitter_light(scenes.lamp_scene(EMITTER_SAMPLING_NONE, False), 1000) == 0.0


def test_an_unseen_emitter_in_the_cell_halves_the_sampled_light():
    # tests.rs:458-469
    want = floor_emitter_expectation() / 2.0
    for strategy in [EMITTER_SAMPLING_ONE, EMITTER_SAMPLING_ALL]:
        got = floor_emitter_light(scenes.lamp_scene(strategy, True), 200_000)
        assert abs((got - want) / want) < 0.02, (strategy, got, want)


def test_an_emitter_seen_from_the_camera_shows_its_emitted_light(oracle):
    # tests.rs:475-492
    s = scenes.lamp_scene(EMITTER_SAMPLING_ONE, False)
    p = s.params
    p["frame"] = [1, 1, 0, 1]
    p["cam_pos"] = [7.5, 2.0, 7.5, 0.01]
    p["cam_fwd"] = [0.0, 1.0, 0.0, 1.0]
    p["cam_right"] = [1.0, 0.0, 0.0, 0.0]
    p["cam_up"] = [0.0, 0.0, -1.0, 0.0]
    c = oracle.render_samples(s.view(), np.zeros(16), np.zeros(16), np.arange(16))
    want = scenes.LAMP_EMITTANCE * scenes.LAMP_INTENSITY
    assert np.all(np.abs(c[:, 0] - want) < 1e-4) and np.all(np.abs(c[:, 1] - want) < 1e-4)


def test_sin_cos_turns_is_within_its_bound_of_f64(oracle):
    # tests.rs:494-520 (MAX_TURN_ERROR = 1.5e-7, math.rs:67)
    inside = (np.arange(1 << 20, dtype=np.float64) / (1 << 20)).astype(np.float32)
    outside = ((np.arange(20000, dtype=np.float32) - np.float32(10000.0)) * np.float32(0.037))
    u = np.concatenate([inside, outside.astype(np.float32)])
    s, c = oracle.sin_cos_turns(u)
    a = 2.0 * np.pi * u.astype(np.float64)
    worst = max(np.abs(s - np.sin(a)).max(), np.abs(c - np.cos(a)).max())
    assert worst <= 1.5e-7, worst
    s, c = oracle.sin_cos_turns(np.array([0.0, 0.25, 0.5, 0.75], "<f4"))
    # Quadrant boundaries are exact, including the signs of zero.
    assert list(s) == [0.0, 1.0, 0.0, -1.0] and list(c) == [1.0, 0.0, -1.0, 0.0]
    assert [bool(np.signbit(x)) for x in s] == [False, False, True, True]
    assert [bool(np.signbit(x)) for x in c] == [False, True, True, False]


def test_ln_and_gamma_decoding_are_within_their_bounds_of_f64(oracle):
    # tests.rs:522-548 (MAX_LN_ERROR 5e-7, MAX_GAMMA_ERROR 1e-6, math.rs:206,210)
    i = np.arange(1 << 20, dtype=np.float32)
    x = (np.float32(1.0 / 255.0) + np.float32(1.0 - 1.0 / 255.0) * i
         / np.float32((1 << 20) - 1)).astype(np.float32)
    worst_ln = np.abs(oracle.ln(x).astype(np.float64) - np.log(x.astype(np.float64))).max()
    c = (np.arange(1, 256, dtype=np.float32) / np.float32(255.0)).astype(np.float32)
    exact = c.astype(np.float64) ** np.float64(np.float32(2.2))
    worst_gamma = np.abs((oracle.gamma_to_linear(c) - exact) / exact).max()
    assert worst_ln <= 5e-7, worst_ln
    assert worst_gamma <= 1e-6, worst_gamma
    g = oracle.gamma_to_linear(np.array([0.0, 1.0], "<f4"))
    assert g[0] == 0.0 and g[1] == 1.0


def test_acos_is_within_its_bound_of_f64(oracle):
    # tests.rs:550-563 (MAX_ACOS_ERROR 1e-6)
    i = np.arange((1 << 20) + 1, dtype=np.float32)
    x = (np.float32(-1.0) + np.float32(2.0) * i / np.float32(1 << 20)).astype(np.float32)
    worst = np.abs(oracle.acos(x) - np.arccos(x.astype(np.float64))).max()
    assert worst <= 1e-6, worst


def test_atan2_is_within_its_bound_of_f64(oracle):
    # tests.rs:565-588 (MAX_ATAN2_ERROR 1e-6)
    a = (np.arange(1 << 16, dtype=np.float32) / np.float32(1 << 16)
         * np.float32(2 * np.pi)).astype(np.float32)
    ys, xs = [], []
    for r in [1e-3, 1.0, 250.0]:
        ys.append(np.sin(a) * np.float32(r))
        xs.append(np.cos(a) * np.float32(r))
    y = np.concatenate(ys).astype(np.float32)
    x = np.concatenate(xs).astype(np.float32)
    worst = np.abs(oracle.atan2(y, x) - np.arctan2(y.astype(np.float64), x.astype(np.float64))).max()
    assert worst <= 1e-6, worst
    got = oracle.atan2(np.array([0.0, 1.0, 0.0], "<f4"), np.array([1.0, 0.0, 0.0], "<f4"))
    for g, w in zip(got, [0.0, np.pi / 2, 0.0]):
        assert abs(float(g) - w) < 1e-7


def test_a_front_face_far_along_the_ray_is_met(oracle):
    # tests.rs:606-683
    v = scenes.wall_scene().view()
    o, d = scenes.wall_rays()
    d = np.array([scenes.normalize32(x) for x in d], "<f4")
    hits = oracle.trace(v, o, d)
    p = o[:, 0] + d[:, 0] * hits["t"]
    ok = (hits["hit"] == 1) & (p > 249.99) & (p < 250.01)
    assert ok.sum() == 4096


def test_a_later_coplanar_quad_is_met(oracle):
    # tests.rs:687-721
    s = scenes.test_scene(False)
    up, north = s.quads[4].copy(), s.quads[0].copy()
    s.quads = np.concatenate([s.quads, [up, north]])
    s.models[1]["quad_count"] = 8
    s.params["off2"][3] = len(s.quads)
    v = s.view()
    for i in range(16):
        f = np.float32(0.1) + np.float32(0.7) * (np.float32(i) + np.float32(0.5)) / np.float32(16.0)
        down = one([np.float32(20.0) + f, 9.0, np.float32(5.0) + f],
                   scenes.normalize32([0.01, -1.0, 0.02]), v, oracle, tmax=100.0)
        assert down is not None and down["quad"] == 6
        south = one([np.float32(20.0) + f, np.float32(5.0) + f, 1.0],
                    scenes.normalize32([0.02, 0.01, 1.0]), v, oracle, tmax=100.0)
        assert south is not None and south["quad"] == 7


def sky_dir(turns, el):
    az = np.float32(turns) * np.float32(2 * np.pi)
    el = np.float32(el)
    return np.array([np.cos(el) * np.cos(az), np.sin(el), np.cos(el) * np.sin(az)], "<f4")


def near(a, b, tol):
    return np.all(np.abs(np.asarray(a) - np.asarray(b)) <= tol)


def test_the_sky_maps_are_read_bilinearly(oracle):
    # tests.rs:756-799
    v = scenes.map_scene().view()
    for j in range(2):
        for i in range(4):
            d = sky_dir((i + 0.5) / 4.0, (j + 0.5) * np.pi / 2 - np.pi / 2)
            assert near(oracle.sky_lens(v, d)[0], [i, j, 1.0], 1e-3)
    assert near(oracle.sky_lens(v, sky_dir(0.0, -np.pi / 4))[0], [1.5, 0.0, 1.0], 1e-3)
    assert near(oracle.sky_lens(v, sky_dir(0.125, 0.0))[0], [0.0, 0.5, 1.0], 1e-3)
    assert near(oracle.sky_lens(v, [0.0, 1.0, 0.0])[0], [1.5, 1.0, 1.0], 1e-3)
    assert near(oracle.sky_lens(v, [0.0, -1.0, 0.0])[0], [1.5, 0.0, 1.0], 1e-3)
    assert near(oracle.sky_light(v, sky_dir(0.25, np.pi / 4))[0], [10.0, 11.0, 2.0], 1e-3)
    assert near(oracle.sky_light(v, sky_dir(0.5, 0.0))[0], [10.5, 10.5, 2.0], 1e-3)


def test_the_lens_sees_one_map_and_a_bounce_is_lit_by_the_other(oracle):
    # tests.rs:804-831
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
    c = oracle.render_samples(s.view(), np.zeros(16), np.zeros(16), np.arange(16))
    assert near(c, [0.3, 0.3, 0.3], 1e-4)
    p["cam_fwd"] = [0.0, 1.0, 0.0, 1.0]
    p["cam_up"] = [0.0, 0.0, 1.0, 0.0]
    assert near(oracle.render_samples(s.view(), [0], [0], [0])[0], [5.0, 5.0, 5.0], 1e-4)
    p["cam_fwd"] = [0.0, -1.0, 0.0, 1.0]
    p["cam_up"] = [0.0, 0.0, -1.0, 0.0]
    p["sun_color"] = [1.0, 1.0, 1.0, 0.03]
    assert oracle.render_samples(s.view(), [0], [0], [0])[0][0] > 1.0
