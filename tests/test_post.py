"""Post-processing (crates/lumenwright/src/render/post.rs) and the PFM writer
(crates/lumenwright-cli/src/output.rs:19-30): the oracle's restatement against the
reference's own tests (post.rs:253-324), and the device method's host half -- the byte
thresholds of the display encoding -- against post.rs's byte for every f32 in [0, 1]."""

import os
import tempfile

import numpy as np

from lumenwright_amd import frames

GAMMA, ACES, FILMIC, MID = 0, 1, 2, 3


def test_tonemap_endpoints(oracle):
    # post.rs:257-266
    px = oracle.tonemap(np.array([[0.0] * 3, [1.0] * 3, [100.0] * 3]), 1.0, GAMMA)
    assert px[0] == 0xFF000000 and px[1] == 0xFFFFFFFF and px[2] == 0xFFFFFFFF
    a = int(oracle.tonemap(np.array([[0.18] * 3]), 1.0, ACES)[0] & 0xFF)
    assert 100 < a < 200
    assert int(oracle.tonemap(np.array([[0.18] * 3]), 1.0, GAMMA)[0] & 0xFF) == 117


def test_the_game_tone_curve_shows_mid_grey_as_the_game(oracle):
    # post.rs:268-302
    curve, k, before = oracle.hable(MID, 0.18)
    assert abs(k - 1.463761) < 1e-6
    assert abs(before - 0.18) < 1e-6
    assert list(curve[:7]) == list(np.array([0.32138956, 0.7318806, 0.1, 0.2, 0.02, 0.3, 7.651521],
                                            dtype=np.float32))
    x = 1e-4
    while x < 7e2:
        _, _, a = oracle.hable(MID, x)
        _, _, b = oracle.hable(FILMIC, k * x)  # the preset at k times the light
        assert abs(a - b) <= 1e-5 * max(abs(b), 1e-6), x
        x *= 1.5
    g = oracle.tonemap(np.array([[0.18] * 3]), 1.0, MID)[0] & 0xFF
    assert g == oracle.tonemap(np.array([[0.18] * 3]), 1.0, GAMMA)[0] & 0xFF
    p = oracle.tonemap(np.array([[0.18] * 3]), 1.0, FILMIC)[0] & 0xFF
    assert p == int(np.float32(0.393285) * np.float32(255.0) + np.float32(0.5))
    w = float(curve[6])
    assert oracle.tonemap(np.array([[w] * 3]), 1.0, MID)[0] == 0xFFFFFFFF


def test_the_tone_mix_is_each_curve_at_its_ends(oracle):
    # post.rs:305-316
    hdr = np.array([[i * 0.05] * 3 for i in range(64)], dtype="<f4")
    assert np.array_equal(oracle.tonemap_mix(hdr, 1.0, GAMMA, MID, 1.0), oracle.tonemap(hdr, 1.0, GAMMA))
    assert np.array_equal(oracle.tonemap_mix(hdr, 1.0, GAMMA, MID, 0.0), oracle.tonemap(hdr, 1.0, MID))
    a, b = oracle.tonemap(hdr, 1.0, GAMMA) & 0xFF, oracle.tonemap(hdr, 1.0, MID) & 0xFF
    m = oracle.tonemap_mix(hdr, 1.0, GAMMA, MID, 0.5) & 0xFF
    assert np.all(m.astype(int) + 1 >= np.minimum(a, b)) and np.all(m <= np.maximum(a, b) + 1)


def test_downsample_averages(oracle):
    # post.rs:319-324
    px = np.array([0xFF000000, 0xFFFFFFFF, 0xFF000000, 0xFFFFFFFF], dtype="<u4")
    assert oracle.downsample(px, 2, 2, 2)[0] & 0xFF == 128


def test_downsample_is_a_rounded_box_mean(oracle):
    # Closed form on random bytes: (sum + n/2) // n per channel.
    rng = np.random.default_rng(3)
    w, h, k = 12, 9, 3
    px = rng.integers(0, 2**32, size=w * h, dtype=np.uint64).astype("<u4")
    got = oracle.downsample(px, w, h, k)
    ch = np.stack([(px >> (8 * i)) & 0xFF for i in range(4)], -1).reshape(h, w, 4).astype(np.int64)
    blocks = ch[: (h // k) * k, : (w // k) * k].reshape(h // k, k, w // k, k, 4).sum(axis=(1, 3))
    want = ((blocks + (k * k) // 2) // (k * k)).astype(np.uint64)
    want = (want[..., 0] | (want[..., 1] << 8) | (want[..., 2] << 16) | (want[..., 3] << 24)).reshape(-1)
    assert np.array_equal(got, want.astype("<u4"))


def test_host_hable_curves_equal_the_oracle(oracle):
    for op in (GAMMA, ACES, FILMIC, MID):
        a = frames.hable_curve(op)
        b, _, _ = oracle.hable(op)
        assert np.array_equal(a.view("<u4"), b.view("<u4")), op


def test_encode_thresholds_give_post_rs_bytes_for_every_f32_in_0_1(oracle):
    thr = frames.encode_thresholds()
    assert np.all(np.diff(thr) >= 0) and thr[0] > 0 and thr[-1] <= 1.0
    bad, first = oracle.check_encode_thresholds(thr)
    assert bad == 0, hex(first)
    # each threshold is the first input of its byte: one f32 below it shows one byte less
    below = (thr.view("<u4") - 1).view("<f4")
    enc_t, enc_b = oracle.encode(thr), oracle.encode(below)
    k = np.arange(1, 256)
    assert np.all((enc_t * np.float32(255.0) + np.float32(0.5)).astype(np.uint32) >= k)
    assert np.all((enc_b * np.float32(255.0) + np.float32(0.5)).astype(np.uint32) < k)


def test_pfm_writer_matches_output_rs(oracle):
    w, h = 5, 3
    hdr = np.random.default_rng(1).random((h, w, 3)).astype("<f4")
    want = oracle.pfm(hdr, w, h)
    assert want.startswith(b"PF\n5 3\n-1.0\n")
    body = np.frombuffer(want[len(b"PF\n5 3\n-1.0\n"):], dtype="<f4").reshape(h, w, 3)
    assert np.array_equal(body, hdr[::-1])  # rows bottom to top
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "x.pfm")
        frames.write_pfm(path, hdr, w, h)
        assert open(path, "rb").read() == want


def test_libm_powf_deviates_only_near_rounding_midpoints(oracle):
    """The device's display encoding takes the power in f64 and rounds it once, and looks up
    libm's powf only for inputs in the band (f64 power within 2^-32 of an f32 rounding
    midpoint). That is exact if every f32 in [0, 1] where powf differs from the rounded power
    is in the band, with room for the f64 power's own error (near 2^-52)."""
    dev, outside, band, not_nb, max_rel = oracle.powf_band()
    assert dev > 0  # libm's powf is not correctly rounded here: the band is needed
    assert outside == 0 and not_nb == 0
    assert max_rel < 2.0 ** -33
    assert band < 1 << 24  # the device sweep's capacity
