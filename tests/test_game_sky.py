"""The game's sky (crates/lumenwright/src/render/sky.rs): the oracle's restatement pinned
against the reference's tests -- the lightmap (sky.rs:1397-1460), the game sky's light
(sky.rs:1462-1563) -- Java's Random, brute-force star counts and closed forms; and the
library's host half (stars, QuickSky::new, the lightmap) against the oracle, bit for bit."""

import ctypes

import numpy as np
import pytest

from lumenwright_amd import frames
from lumenwright_amd.abi import BoundView, new_params


def test_java_random_known_answers(oracle):
    # java.util.Random: new Random(42).nextInt() and new Random(0).nextInt().
    assert oracle.java_random(42, 32, 1)[0] == -1170105035
    assert oracle.java_random(0, 32, 1)[0] == -1155484576


def lightmap_fsh(sky_factor, colour, brightness):
    """lightmap.fsh of the 1.21.11 client in f64 at sky light 15, block light 0 (sky.rs:1397-1425)."""
    def get_brightness(level):
        return level / (4.0 - 3.0 * level)
    block = get_brightness(0.0) * 1.5
    sky = get_brightness(15.0 / 15.0) * sky_factor
    c = [block, block * ((block * 0.6 + 0.4) * 0.6 + 0.4), block * (block * block * 0.6 + 0.4)]
    for k in range(3):
        c[k] += colour[k] * sky
        c[k] = c[k] * 0.96 + 0.75 * 0.04
        c[k] = min(max(c[k], 0.0), 1.0)
    mx = max(c)
    inv = 1.0 - mx
    scaled = 1.0 - inv * inv * inv * inv
    out = []
    for k in range(3):
        ng = c[k] * (scaled / mx)
        v = c[k] + (ng - c[k]) * brightness
        out.append(v * 0.96 + 0.75 * 0.04)
    return out


def test_the_sky_light_is_the_clients_lightmap(oracle):
    # sky.rs:1430-1459
    checked = 0
    for f in (0.0, 0.24, 0.385497, 0.62, 1.0):
        for col in ([1.0, 1.0, 1.0], [0x7A / 255.0, 0x7A / 255.0, 1.0]):
            col32 = np.array(col, "<f4")
            for b in (0.0, 0.5, 1.0):
                got = oracle.lightmap_sky(f, col32, b)
                assert np.array_equal(got, frames.lightmap_sky(f, col32, b))
                want = lightmap_fsh(float(np.float32(f)), [float(c) for c in col32], float(np.float32(b)))
                for k in range(3):
                    assert abs(float(got[k]) - want[k]) <= 0.5 / 255 + 1e-6
                    assert float(got[k] * np.float32(255.0)) % 1.0 == 0.0  # an 8-bit level
                checked += 1
    assert checked == 30
    c = np.float32(0x93) / np.float32(255.0)
    light = oracle.lightmap_sky(0.385497, np.array([c, c, 1.0], "<f4"), 0.5).astype(np.float64) ** 2.2
    engine = 0.374446
    level = round(engine ** (1 / 2.2) * 255) / 255
    assert abs(light[2] - level ** 2.2) < 2e-3


def test_the_stars_are_the_clients(oracle):
    s = oracle.stars()
    assert abs(len(s) - 1500 * np.pi / 6) < 40  # the draws inside the unit ball of [-1, 1]^3
    c, a1, a2 = s[:, :3].astype(np.float64), s[:, 3:6].astype(np.float64), s[:, 6:9].astype(np.float64)
    assert np.allclose(np.linalg.norm(c, axis=1), 100.0, rtol=1e-5)
    size = np.linalg.norm(a1, axis=1)
    assert np.allclose(size, np.linalg.norm(a2, axis=1), rtol=1e-5)
    assert size.min() >= 0.15 - 1e-6 and size.max() <= 0.25 + 1e-6
    for u, v in ((a1, c), (a2, c), (a1, a2)):
        cosang = np.abs(np.sum(u * v, axis=1)) / (np.linalg.norm(u, axis=1) * np.linalg.norm(v, axis=1))
        assert cosang.max() < 1e-5
    assert np.array_equal(oracle.star_table(), frames.star_table())


def dusk_description(with_bodies: bool):
    """sky.rs:1462-1481: the dusk sky at tick 12800, the camera facing the sun's side."""
    sun = frames.celestial_direction(90.29)
    return frames.sky_description(
        sky=(0.18, 0.24, 0.37), fog=(0.62, 0.40, 0.33), sky_fog_end=192.0,
        render_fog=(172.8, 192.0), env_fog=(0.0, 1024.0), sky_disc=True,
        sunrise=(0.98, 0.57, 0.25, 0.62), sun_angle=90.29,
        sun=(sun, (0.0, 0.0, 2.0, 2.0)) if with_bodies else None, moon=None, star_angle=90.29,
        star_brightness=0.38 if with_bodies else 0.0)


def dusk_camera():
    """sky.rs:1483-1491"""
    def unit(v):
        v = np.array(v, "<f4")
        return v * (np.float32(1.0) / np.sqrt(np.sum(v * v, dtype=np.float32)))
    cam = frames.CameraRaysC()
    cam.forward[:] = [float(x) for x in unit([-0.83, -0.35, 0.43])]
    cam.right[:] = [float(x) for x in unit([-0.46, 0.0, -0.89])]
    cam.up[:] = [0.0, 1.0, 0.0]
    cam.tan_half_fov, cam.aspect = 0.577, 16.0 / 9.0
    return cam


def test_quick_sky_new_and_celestial_directions_are_the_oracles(oracle):
    for deg in (0.0, 45.0, 90.29, 180.0, 271.5):
        a = np.radians(deg)
        got = frames.celestial_direction(deg)
        assert np.array_equal(got, oracle.celestial_direction(deg))
        assert np.allclose(got, [-np.sin(a), np.cos(a), 0.0], atol=1e-6)  # Y about z then y
    cam = dusk_camera()
    for bodies in (True, False):
        d = dusk_description(bodies)
        a = bytes(frames.quick_sky_new(cam, 16, 9, d))
        b = bytes(oracle.quick_sky_new(cam, 16, 9, d))
        assert a == b


def star_sky(brightness=0.1):
    """A black sky with stars only: each star along a direction adds brightness^2."""
    return frames.sky_description(
        sky=(0, 0, 0), fog=(0, 0, 0), sky_fog_end=192.0, render_fog=(172.8, 192.0),
        env_fog=(0.0, 1024.0), sky_disc=True, sunrise=(0, 0, 0, 0), sun_angle=90.0,
        star_angle=33.0, star_brightness=brightness)


def test_every_star_along_a_direction_is_counted(oracle):
    # The binned lookup (sky.rs:486-507) against every star: directions at and around the
    # stars' own centres and random ones.
    q = oracle.quick_sky_new(dusk_camera(), 16, 9, star_sky())
    s = oracle.stars().astype("<f4")
    rot = np.array([list(q.star_rot[i])[:3] for i in range(3)], "<f4")
    rng = np.random.default_rng(5)
    dirs = []
    for st in s[rng.choice(len(s), 300, replace=False)]:
        c = st[:3] / np.float32(100.0)
        for k in range(6):
            jitter = rng.normal(size=3) * (0.0 if k == 0 else 0.0015 * k)
            q_dir = c + jitter
            dirs.append(rot.T.astype(np.float64) @ (q_dir / np.linalg.norm(q_dir)))
    dirs += list(rng.normal(size=(2000, 3)))
    d = np.array([v / np.linalg.norm(v) for v in dirs], "<f4")
    got = oracle.quick_background(q, d, np.array([0xFFFFFFFF], "<u4"), 1)[:, 0]
    n_got = np.rint(got / np.float32(0.01)).astype(int)
    qd = (d.astype("<f4") @ rot.T).astype("<f4")
    a1 = s[:, 3:6] / np.sum(s[:, 3:6] * s[:, 3:6], axis=1, keepdims=True)
    a2 = s[:, 6:9] / np.sum(s[:, 6:9] * s[:, 6:9], axis=1, keepdims=True)
    cc = s[:, :3]
    dn = qd @ cc.T
    with np.errstate(divide="ignore", invalid="ignore"):
        p = qd[:, None, :] * (np.sum(cc * cc, axis=1) / dn)[:, :, None] - cc[None, :, :]
        inside = (dn > 0) & (np.abs(np.sum(p * a1[None], axis=2)) < 1) & (np.abs(np.sum(p * a2[None], axis=2)) < 1)
    n_want = inside.sum(axis=1)
    assert (n_want > 0).sum() > 200
    assert np.array_equal(n_got, np.minimum(n_want, 100))


def test_the_game_sky_casts_the_lightmaps_light(oracle):
    # sky.rs:1492-1563
    atlas = np.full(4, 0xFFFFFFFF, "<u4")
    cam = dusk_camera()
    lens = oracle.quick_sky_new(cam, 16, 9, dusk_description(True))
    bare = oracle.quick_sky_new(cam, 16, 9, dusk_description(False))
    colour = np.array([0x92, 0x92, 0xFF], "<f4") / np.float32(255.0)
    maps, info = oracle.game_sky_maps(lens, bare, 0.378182, colour, atlas, 2)
    assert maps.size == (2048 * 1024 + 360 * 180) * 3
    veil, scale = np.array(info["veil"]), info["scale"]
    assert veil.min() >= 0.0 and (veil.min() == 0.0 or scale == 50.0)
    p = new_params()
    p["sky_map"] = [2048, 1024, 360, 180]
    z = np.zeros(1, "<u4")
    v = BoundView(p, z, np.zeros(1, dtype=frames.MODEL), np.zeros(1, dtype=frames.QUAD), z,
                  np.zeros(1, dtype=frames.EMITTER), np.zeros(8, "<u4"), maps)
    na, ne = 720, 180
    az = (np.arange(na) + 0.5) / na * 2 * np.pi
    el = (np.arange(ne) + 0.5) / ne * np.pi / 2
    A, E = np.meshgrid(az, el, indexing="ij")
    d = np.stack([np.cos(E) * np.cos(A), np.sin(E), np.cos(E) * np.sin(A)], -1).reshape(-1, 3)
    c = oracle.sky_light(v, d.astype("<f4")).astype(np.float64)
    w = (np.sin(E) * np.cos(E)).reshape(-1)
    cast = (c * w[:, None]).sum(0) / w.sum()
    light = np.array(info["light"], np.float64)
    assert np.all(np.abs(cast - light) / light < 2e-3), (cast, light)
    sun = frames.celestial_direction(90.29)
    ds = np.array([[sun[0], max(sun[1], np.float32(0.01)), sun[2]]], "<f4")
    seen = oracle.sky_lens(v, ds)[0]
    lit = oracle.sky_light(v, ds)[0] / np.float32(scale)
    assert seen[2] > 0.9 and seen[2] > lit[2] + 0.3, (seen, lit)


def test_libm_powf_band_holds_for_the_texel_decoding(oracle):
    # The device's c^2.2 (pow_libm): every f32 in [0, 1] where libm's powf(x, 2.2) differs
    # from the rounded f64 power lies in the band, with margin.
    dev, outside, band, not_nb, max_rel = oracle.powf_band(exponent=2.2)
    assert dev > 0 and outside == 0 and not_nb == 0
    assert max_rel < 0.75 * 2.0 ** -32  # 2^-32.9: a margin far above the f64 power's 2^-52
    assert band < 1 << 24
