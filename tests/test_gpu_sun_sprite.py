"""Chunky's sun sprite on the device (sun_texture, crates/lumenwright-kernel/src/integrator.rs:
249-277, added to camera rays that escape, :620): no configuration draws it, so here the
config-0 frame looks at its sun with a sprite from the block atlas, and the device's samples
equal the oracle's bit for bit."""

import numpy as np
import pytest

from lumenwright_amd import frames

pytestmark = pytest.mark.gpu


def sun_frame(sprite):
    sun = frames.sun_direction_from_angle(frames.celestial_angle(4000)).astype(np.float64)
    yaw = float(np.degrees(np.arctan2(-sun[0], sun[2])))
    pitch = float(np.degrees(np.arcsin(-sun[1])))
    f = frames.make_frame(frames.config(0, yaw=yaw, pitch=pitch, width=160, height=90))
    f.params["sun_sprite"] = sprite
    return f


def test_the_sun_sprite_on_the_device_is_the_oracles(hip_backend, oracle):
    f = sun_frame([0, 0, 16, 16])
    v = f.view()
    rng = np.random.default_rng(2)
    px = rng.integers(60, 100, 20000)
    py = rng.integers(25, 65, 20000)
    s = rng.integers(0, 16, 20000)
    want = oracle.render_samples(v, px, py, s)
    got = hip_backend.render_samples(v, px, py, s)
    bad = np.flatnonzero(np.any(got.view("<u4") != want.view("<u4"), axis=1))
    assert bad.size == 0, (bad.size, got[bad[:3]], want[bad[:3]])
    # the sprite is drawn: the same samples without it differ where the camera sees the sun
    plain = sun_frame([0, 0, 0, 0])
    base = oracle.render_samples(plain.view(), px, py, s)
    drawn = np.any(base != want, axis=1)
    assert 200 < drawn.sum() < 20000
