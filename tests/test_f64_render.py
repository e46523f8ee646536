"""The f64 render (oracle/lw_oracle_f64.cpp): the path tracer with every operation in double and
libm's double functions, against which the f32 path's own rounding error is stated. The f32
path here is the reference's CPU backend (its op-for-op restatement), and the device's image
equals it bit for bit, so the error is the device's too.

Pins of the f64 render itself: the reference's known answers that are exact in any precision
(the lamp's emission 0.5 * 13, tests.rs:475-492; the lens and light maps 0.3, tests.rs:804-831),
and agreement with the f32 render to f32 rounding on nearly every sample."""

import numpy as np

import scenes
from lumenwright_amd import frames
from lumenwright_amd.abi import EMITTER_SAMPLING_ONE


def test_known_answers_in_f64(oracle):
    s = scenes.lamp_scene(EMITTER_SAMPLING_ONE, False)
    p = s.params
    p["frame"] = [1, 1, 0, 1]
    p["cam_pos"] = [7.5, 2.0, 7.5, 0.01]
    p["cam_fwd"] = [0.0, 1.0, 0.0, 1.0]
    p["cam_right"] = [1.0, 0.0, 0.0, 0.0]
    p["cam_up"] = [0.0, 0.0, -1.0, 0.0]
    c = oracle.render_samples_f64(s.view(), np.zeros(16), np.zeros(16), np.arange(16))
    assert np.all(np.abs(c[:, :2] - 6.5) < 1e-12)
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
    c = oracle.render_samples_f64(s.view(), np.zeros(16), np.zeros(16), np.arange(16))
    assert np.all(np.abs(c - np.float32(0.3)) < 1e-12)


def f32_against_f64(oracle, frame, spp):
    v = frame.view()
    a = oracle.render_hq(v, spp).astype(np.float64)
    b = oracle.render_hq_f64(v, spp)
    return np.abs(a - b), b


def test_the_f32_images_error_against_the_f64_render(oracle):
    # config 0 in full (320x180, 16 spp) and config 1's world at 320x180, 4 spp
    for name, frame, spp in [("cfg0", frames.make_frame(frames.CONFIGS[0]), 16),
                             ("cfg1", frames.make_frame(frames.config(1, width=320, height=180, spp=4)), 4)]:
        d, ref = f32_against_f64(oracle, frame, spp)
        mae = d.mean()
        print(f"{name}: MAE {mae:.3e} (relative {mae / ref.mean():.2e}), "
              f"pixels off by more than 1e-3: {(d.max(axis=-1) > 1e-3).mean():.4%}")
        assert mae < 1e-3
        assert mae / ref.mean() < 2e-4


def test_samples_agree_to_f32_rounding(oracle):
    # A sample agrees within 1e-4 relative unless a comparison on its path landed the other
    # way in f64 (a shadow edge, a Russian-roulette or Fresnel draw at its threshold).
    for frame, seed in [(frames.make_frame(frames.CONFIGS[0]), 1),
                        (frames.make_frame(frames.config(1, width=320, height=180, spp=16)), 2)]:
        v = frame.view()
        w, h = int(frame.params["frame"][0]), int(frame.params["frame"][1])
        rng = np.random.default_rng(seed)
        px, py, s = rng.integers(0, w, 20000), rng.integers(0, h, 20000), rng.integers(0, 16, 20000)
        x32 = oracle.render_samples(v, px, py, s).astype(np.float64)
        x64 = oracle.render_samples_f64(v, px, py, s)
        rel = np.max(np.abs(x32 - x64) / np.maximum(np.abs(x64), 1e-3), axis=1)
        assert (rel <= 1e-4).mean() > 0.99


def test_the_f64_row_render_is_the_f64_frames_rows(oracle):
    # bench.py states the 720p frame's error against f64 on a subset of rows
    from lumenwright_amd import frames
    fr = frames.make_frame(frames.config(1, width=80, height=45, world_chunks_x=4, world_chunks_z=4))
    rows = [0, 7, 22, 44]
    whole = oracle.render_hq_f64(fr.view(), 3)
    assert np.array_equal(oracle.render_rows_f64(fr.view(), 3, rows), whole[rows])
