"""Post-processing on the device (lw_hip_tonemap, lw_hip_tonemap_mix, lw_hip_downsample,
lw_hip_render_hq_rgba8) against the oracle's restatement of crates/lumenwright/src/render/
post.rs: the packed RGBA8 bytes must be equal, pixel for pixel."""

import numpy as np
import pytest

from lumenwright_amd import frames

pytestmark = pytest.mark.gpu

OPS = (frames.TONEMAP_GAMMA, frames.TONEMAP_ACES, frames.TONEMAP_FILMIC, frames.TONEMAP_FILMIC_MID_GREY)


def hdr_cases(n_random: int = 1 << 20, seed: int = 7) -> np.ndarray:
    """Random radiance over seven decades, then the edges: signed zeros, negatives,
    subnormals, NaN, infinities, 1 and every display threshold with its two neighbours."""
    rng = np.random.default_rng(seed)
    r = np.exp(rng.uniform(np.log(1e-5), np.log(1e3), size=n_random * 3)).astype("<f4")
    thr = frames.encode_thresholds()
    bits = thr.view("<u4")
    near = np.concatenate([thr, (bits - 1).view("<f4"), (bits + 1).view("<f4")])
    edges = np.array([0.0, -0.0, -1.0, 1.0, 100.0, np.nan, np.inf, -np.inf, 1e-40, 1e-45,
                      np.float32(3.4e38), 0.18, 0.5, 7.651521, 11.2], dtype="<f4")
    v = np.concatenate([r, near, edges])
    v = np.concatenate([v, np.zeros((-v.size) % 3, dtype="<f4")])
    return v.reshape(-1, 3)


@pytest.fixture(scope="module")
def hdr():
    return hdr_cases()


@pytest.mark.parametrize("op", OPS)
@pytest.mark.parametrize("exposure", [1.0, 0.37, 4.0])
def test_tonemap_equals_the_oracle(hip_backend, oracle, hdr, op, exposure):
    got = hip_backend.tonemap(hdr, exposure, op)
    want = oracle.tonemap(hdr, exposure, op)
    bad = np.flatnonzero(got != want)
    assert bad.size == 0, (bad.size, hdr[bad[:4]], got[bad[:4]], want[bad[:4]])


def test_tonemap_at_the_edges_of_the_image(hip_backend, oracle):
    for n in (0, 1, 255, 256, 257):
        h = np.linspace(0.0, 2.0, n * 3, dtype="<f4").reshape(-1, 3)
        assert np.array_equal(hip_backend.tonemap(h, 1.0, 0), oracle.tonemap(h, 1.0, 0))


@pytest.mark.parametrize("pair", [(0, 3), (1, 2), (3, 3)], ids=["gamma-game", "aces-filmic", "game-game"])
@pytest.mark.parametrize("weight", [0.0, 1.0, 0.5, 0.25, 0.9, 1e-3, -2.0, 3.0])
def test_tonemap_mix_equals_the_oracle(hip_backend, oracle, hdr, pair, weight):
    a, b = pair
    got = hip_backend.tonemap_mix(hdr, 1.3, a, b, weight)
    want = oracle.tonemap_mix(hdr, 1.3, a, b, weight)
    bad = np.flatnonzero(got != want)
    assert bad.size == 0, (bad.size, hdr[bad[:4]], got[bad[:4]], want[bad[:4]])


def test_display_encoding_equals_libm_powf(hip_backend, oracle):
    """lw_hip_encode against libm's powf on every input of the band (where the rounded f64
    power and powf can differ), 8M random inputs of [0, 1] and the edges."""
    *_, band = oracle.powf_band(inputs=True)
    rng = np.random.default_rng(11)
    rand = rng.integers(0, 0x3F800001, size=1 << 23, dtype=np.uint64).astype("<u4")
    edges = np.array([0.0, -0.0, 1.0, 1.5, -2.0, np.inf, -np.inf, np.nan, 1e-45, 1e-38],
                     dtype="<f4")
    x = np.concatenate([band.view("<f4"), rand.view("<f4"), edges])
    got = hip_backend.encode(x)
    print(f"band inputs {band.size}, table set-up {hip_backend.encode_setup_ms:.0f} ms")
    want = oracle.encode(x)
    nan = np.isnan(want)
    assert np.array_equal(np.isnan(got), nan)
    bad = np.flatnonzero(got[~nan].view("<u4") != want[~nan].view("<u4"))
    assert bad.size == 0, (bad.size, x[~nan][bad[:4]], got[~nan][bad[:4]], want[~nan][bad[:4]])


@pytest.mark.parametrize("w,h,k", [(2, 2, 2), (1280, 720, 2), (1283, 721, 3), (97, 13, 4),
                                   (64, 64, 1), (5, 3, 8), (1920, 1080, 5)])
def test_downsample_equals_the_oracle(hip_backend, oracle, w, h, k):
    px = np.random.default_rng(w * 7 + h + k).integers(0, 2**32, size=w * h, dtype=np.uint64)
    px = px.astype("<u4")
    got = hip_backend.downsample(px, w, h, k)
    assert np.array_equal(got, oracle.downsample(px, w, h, k))


@pytest.mark.parametrize("op", OPS)
def test_render_then_tonemap_on_the_device_equals_the_oracle_tonemap(hip_backend, oracle, op):
    fr = frames.make_frame(frames.config(1, width=160, height=90, spp=4,
                                         world_chunks_x=6, world_chunks_z=6))
    scene = hip_backend.upload(fr)
    try:
        params = fr.view().params
        hdr = hip_backend.render_hq_scene(scene, params, fr.spp)
        px = hip_backend.render_hq_rgba8(scene, params, fr.spp, 0.8, op)
        assert hip_backend.last_timing()["samples"] == 160 * 90 * 4
        assert np.array_equal(px, oracle.tonemap(hdr, 0.8, op))
        assert np.array_equal(px, hip_backend.tonemap(hdr, 0.8, op))
    finally:
        scene.free()


def test_unknown_operator_is_an_error(hip_backend):
    import lumenwright_amd as lw
    with pytest.raises(lw.BackendError):
        hip_backend.tonemap(np.zeros((4, 3), dtype="<f4"), 1.0, 9)
    with pytest.raises(lw.BackendError):
        hip_backend.tonemap_mix(np.zeros((4, 3), dtype="<f4"), 1.0, 0, 9, 0.5)
