"""The host-side frame preparation (include/lw/host.h) against the reference's own tests:
crates/lumenwright/src/render/sky.rs:954-1397 (the sun's day cycle, twilight, Preetham
against an f64 transcription of Chunky, the sky cache against Chunky's SkyCache),
crates/lumenwright/src/scene/camera.rs:86-131 and the emitter grid guarantee of
crates/lumenwright/src/scene/voxel.rs:616-671. The f64 transcriptions below restate
Chunky's published formulas (A. J. Preetham, P. Shirley, B. Smits, SIGGRAPH 1999; Chunky's
SkyCache bilinear HSL lookup) as the reference's tests do."""

import numpy as np
import pytest

import scenes
from lumenwright_amd import frames
from lumenwright_amd.abi import BoundView, new_params

SUN_RADIUS = np.float32(0.03)
CIVIL_TWILIGHT_END = np.float32(-6.0) * np.float32(np.pi) / np.float32(180.0)
NIGHT_SKY_LIGHT = np.float32(0.03)


def test_sun_follows_vanilla_day_cycle():
    # sky.rs:958-976
    assert abs(frames.celestial_angle(6000)) < 1e-4
    assert frames.sun_direction_from_angle(frames.celestial_angle(6000))[1] > 0.999
    m = frames.sun_direction_from_angle(frames.celestial_angle(0))
    assert m[0] > 0.95 and 0.1 < m[1] < 0.3
    e = frames.sun_direction_from_angle(frames.celestial_angle(12000))
    assert e[0] < -0.95 and e[1] > 0.0


def test_twilight_holds_chunkys_sky_until_the_disc_sets():
    # sky.rs:1112-1131
    tw = frames.twilight
    assert tw(0.5) == 1.0 and tw(0.0) == 1.0 and tw(-0.03) == 1.0
    assert tw(-0.03 - 1e-3) < 1.0
    assert tw(float(CIVIL_TWILIGHT_END)) == np.float32(0.03)
    assert tw(-np.pi / 2) == np.float32(0.03)
    last = 1.0
    for i in range(101):
        t = tw(float(-SUN_RADIUS + (CIVIL_TWILIGHT_END + SUN_RADIUS) * np.float32(i) / np.float32(100)))
        assert t <= last
        last = t
    dusk = frames.hq_sky(frames.sun_direction(float(np.degrees(np.float32(-0.005091))), 270.0), 1.25, 1.0)
    assert dusk.sky_light == 1.0
    assert list(dusk.sun_color) == [np.float32(np.float32(1.25) ** np.float32(2.2))] * 3
    assert frames.hq_sky([0, -1, 0], 1.25, 1.0).sky_light == np.float32(0.03)


def test_sun_texture_axes():
    # sky.rs:1134-1150
    su, sv = frames.sun_texture_axes([-0.999987, -0.005091, 0.0])
    assert np.linalg.norm(su - [-0.005091, 0.999987, 0.0]) < 1e-5
    assert np.linalg.norm(sv - [0.0, 0.0, -1.0]) < 1e-5


def test_camera_conventions():
    # camera.rs:86-131
    def fwd(y, p):
        return frames.camera_basis(y, p)[0]
    assert np.linalg.norm(fwd(0, 0) - [0, 0, 1]) < 1e-5
    assert np.linalg.norm(fwd(90, 0) + [1, 0, 0]) < 1e-5
    assert np.linalg.norm(fwd(180, 0) + [0, 0, 1]) < 1e-5
    assert np.linalg.norm(fwd(0, 90) + [0, 1, 0]) < 1e-5
    f, r, u = frames.camera_basis(30, 20)
    assert u[1] > 0 and np.linalg.norm(np.cross(r, u) + f) < 1e-5
    _, r, _ = frames.camera_basis(0, 0)
    assert np.linalg.norm(r + [1, 0, 0]) < 1e-5


def test_preetham_matches_the_reference_constants():
    # tests.rs:590-600: Preetham::new for a sun at altitude 45, azimuth 45 degrees.
    sun = np.array([0.5, np.sqrt(0.5), 0.5], "<f4")
    sky = frames.preetham_new(sun, 2.5)
    np.testing.assert_allclose(list(sky.zenith), scenes.PREETHAM_45_ZENITH[:3], rtol=1e-6)
    for i in range(5):
        assert list(sky.perez[i]) == [np.float32(x) for x in scenes.PREETHAM_PEREZ[i][:3]]


# ---- Chunky's PreethamSky.calcIncidentLight in f64 (the reference's test transcription) ----
MDX = np.array([[-0.0193, -0.2592], [-0.0665, 0.0008], [-0.0004, 0.2125], [-0.0641, -0.8989],
                [-0.0033, 0.0452]])
MDY = np.array([[-0.0167, -0.2608], [-0.0950, 0.0092], [-0.0079, 0.2102], [-0.0441, -1.6537],
                [-0.0109, 0.0529]])
MDL = np.array([[0.1787, -1.4630], [-0.3554, 0.4275], [-0.0227, 5.3251], [0.1206, -2.5771],
                [-0.0670, 0.3703]])
XZ = np.array([[0.00166, -0.00375, 0.00209, 0.0], [-0.02903, 0.06377, -0.03203, 0.00394],
               [0.11693, -0.21196, 0.06052, 0.25886]])
YZ = np.array([[0.00275, -0.00610, 0.00317, 0.0], [-0.04214, 0.08970, -0.04153, 0.00516],
               [0.15346, -0.26756, 0.06670, 0.26688]])


def chunky_preetham(alt, az, d, turb=2.5):
    """Preetham sky radiance (linear RGB, x0.045) for directions d (n, 3), in f64."""
    d = np.asarray(d, dtype=np.float64)
    a, b, cl = (m[:, 0] * turb + m[:, 1] for m in (MDX, MDY, MDL))

    def perez(ct, g, c2g, k):
        with np.errstate(divide="ignore"):
            first = 1.0 + k[0] * np.exp(k[1] / ct)
        return first * (1.0 + k[2] * np.exp(k[3] * g) + k[4] * c2g)

    def chroma(st, m):
        r = lambda row: row[0] * st ** 3 + row[1] * st ** 2 + row[2] * st + row[3]
        return turb * turb * r(m[0]) + turb * r(m[1]) + r(m[2])

    st = np.pi / 2 - alt
    c2t = np.cos(st) ** 2
    chi = (4.0 / 9.0 - turb / 120.0) * (np.pi - 2.0 * st)
    zy = abs((4.0453 * turb - 4.9710) * np.tan(chi) - 0.2155 * turb + 2.4192)
    zx, zyc = chroma(st, XZ), chroma(st, YZ)
    f0x, f0y, f0l = (1.0 / perez(1.0, st, c2t, k) for k in (a, b, cl))
    r = abs(np.cos(alt))
    sw = np.array([np.cos(az) * r, np.sin(alt), np.sin(az) * r])
    ct = np.maximum(d[:, 1], 0.0)
    cg = d @ sw
    g = np.arccos(np.clip(cg, -1, 1))
    x = zx * perez(ct, g, cg * cg, a) * f0x
    y = zyc * perez(ct, g, cg * cg, b) * f0y
    z = zy * perez(ct, g, cg * cg, cl) * f0l
    f = z / y
    x2, y2, z2 = x * f, z, (1.0 - x - y) * f
    return np.stack([(2.3706743 * x2 - 0.9000405 * y2 - 0.4706338 * z2) * 0.045,
                     (-0.513885 * x2 + 1.4253036 * y2 + 0.0885814 * z2) * 0.045,
                     (0.0052982 * x2 - 0.0146949 * y2 + 1.0093968 * z2) * 0.045], 1)


def sky_params(sun, turb=2.5, res=0.0):
    sky = frames.preetham_new(sun, turb)
    p = new_params()
    p["sky_sun"] = list(sky.sun) + [1.0]
    p["sky_zenith"] = list(sky.zenith) + [0.0]
    for k, name in enumerate(["perez_a", "perez_b", "perez_c", "perez_d", "perez_e"]):
        p[name] = list(sky.perez[k]) + [0.0]
    p["misc_f"][2] = res
    return p


def empty_view(p, sky=()):
    return BoundView(p, np.zeros(1, "<u4"), np.zeros(1, frames.MODEL), np.zeros(1, frames.QUAD),
                     np.zeros(1, "<u4"), np.zeros(1, frames.EMITTER), np.zeros(8, "<u4"),
                     np.asarray(sky, "<f4"))


@pytest.mark.parametrize("turb", [2.5, 3.5])
def test_the_kernel_sky_is_chunkys_preetham_sky(oracle, turb):
    # sky.rs:1053-1108: preetham_radiance (the oracle's) fed from Preetham::new (ours) agrees
    # with Chunky's f64 formula within 1e-4 relative; turbidity 3.5 is the inverse-render
    # target the reference does not state.
    worst, n = 0.0, 0
    for alt, az in [(45, 45), (10, 200), (80, 300), (2, 90)]:
        alt, az = np.radians(alt), np.radians(az)
        sun = np.array([np.cos(az) * np.cos(alt), np.sin(alt), np.sin(az) * np.cos(alt)], "<f4")
        v = empty_view(sky_params(sun, turb))
        i, j = np.meshgrid(np.arange(64), np.arange(17), indexing="ij")
        phi = i.ravel() / 64.0 * 2 * np.pi
        h = j.ravel() / 16.0 * 1.2 - 0.2
        r = np.sqrt(1.0 - h * h)
        d = np.stack([np.cos(phi) * r, h, np.sin(phi) * r], 1)
        want = chunky_preetham(alt, az, d, turb)
        got = oracle.preetham_radiance(v, d.astype("<f4"))
        worst = max(worst, float(np.max(np.abs(got - want) / np.maximum(np.abs(want), 0.01))))
        n += len(d)
    assert n == 4 * 64 * 17
    assert worst < 1e-4, worst


def to_hsl(c):
    c = np.clip(c, 0, 1)
    r, g, b = c[..., 0], c[..., 1], c[..., 2]
    cmax, cmin = c.max(-1), c.min(-1)
    delta = cmax - cmin
    l = (cmax + cmin) / 2
    with np.errstate(divide="ignore", invalid="ignore"):
        h = np.where(delta == 0, 0.0,
                     np.where(cmax == r, np.fmod((g - b) / delta, 6.0) / 6.0,
                              np.where(cmax == g, ((b - r) / delta + 2.0) / 6.0,
                                       ((r - g) / delta + 4.0) / 6.0)))
        s = np.where(delta < 5e-8, 0.0, delta / (1.0 - np.abs(2 * l - 1)))
    return np.stack([h, s, l], -1)


def from_hsl(hsl):
    h, s, l = hsl[..., 0], hsl[..., 1], hsl[..., 2]
    c = np.minimum((1 - np.abs(2 * l - 1)) * s, 1.0)
    h6 = h * 6
    x = c * (1 - np.abs(np.fmod(h6, 2.0) - 1))
    z = np.zeros_like(c)
    hh = h6[..., None]
    rgb = np.select([hh < 1, hh < 2, hh < 3, hh < 4, hh < 5],
                    [np.stack(t, -1) for t in [(c, x, z), (x, c, z), (z, c, x), (z, x, c), (x, z, c)]],
                    np.stack([c, z, x], -1))
    m = np.maximum(l - 0.5 * c, 0.0)
    return rgb + m[..., None]


def test_the_kernel_reads_the_sky_through_chunkys_sky_cache(oracle):
    # sky.rs:1255-1397: sky_radiance (the oracle's) reading our sky cache agrees with
    # Chunky's SkyCache over its f64 table within 1e-4 absolute.
    n = 128
    worst, count, saturated = 0.0, 0, 0
    for alt, az in [(-0.005091, np.pi), (0.1, 2.0), (0.8, 0.7)]:
        ii, jj = np.meshgrid(np.arange(n + 1), np.arange(n + 1), indexing="ij")
        theta = ii / n * 2 * np.pi
        phi = jj / n * np.pi - np.pi / 2
        dirs = np.stack([np.cos(theta) * np.cos(phi), np.sin(phi), np.sin(theta) * np.cos(phi)], -1)
        table = to_hsl(chunky_preetham(max(alt, 0.0), az, dirs.reshape(-1, 3))).reshape(n + 1, n + 1, 3)
        sun = np.array([np.cos(az) * np.cos(alt), np.sin(alt), np.sin(az) * np.cos(alt)], "<f4")
        p = sky_params(sun, 2.5, float(n))
        cache = frames.sky_cache(p)
        assert cache.size == (n + 1) ** 2 * 3
        v = empty_view(p, cache)
        i, j = np.meshgrid(np.arange(97), np.arange(41), indexing="ij")
        a = i.ravel() / 97.0 * 2 * np.pi + 0.013
        h = j.ravel() / 40.0 * 1.9 - 0.95
        r = np.sqrt(1.0 - h * h)
        d = np.stack([np.cos(a) * r, h, np.sin(a) * r], 1)
        th = np.arctan2(d[:, 2], d[:, 0]) / (2 * np.pi)
        th = np.fmod(np.fmod(th, 1.0) + 1.0, 1.0)
        ph = (np.arcsin(np.clip(d[:, 1], -1, 1)) + np.pi / 2) / np.pi
        x, y = th * n, ph * n
        fx = np.clip(x, 0, n - 1).astype(int)
        fy = np.clip(y, 0, n - 1).astype(int)
        y0 = table[fx, fy] + (x - fx)[:, None] * (table[fx + 1, fy] - table[fx, fy])
        y1 = table[fx, fy + 1] + (x - fx)[:, None] * (table[fx + 1, fy + 1] - table[fx, fy + 1])
        want = from_hsl(y0 + (y - fy)[:, None] * (y1 - y0))
        got = oracle.sky_radiance(v, d.astype("<f4"))
        worst = max(worst, float(np.abs(got - want).max()))
        direct = chunky_preetham(max(alt, 0.0), az, d)
        sat = np.any(direct > 1.0, axis=1)
        saturated += int(sat.sum())
        assert np.all(got[sat] <= 1.0 + 1e-5)
        count += len(d)
    assert count == 3 * 97 * 41 and saturated > 0
    assert worst < 1e-4, worst


def test_our_sky_cache_and_gamma_table_equal_the_oracles(oracle):
    # The product builds the sky cache (integrator.rs:133-148) and the gamma table on the
    # host; both must be bit-identical to the restated kernel's values.
    sun = frames.sun_direction_from_angle(frames.celestial_angle(4000))
    p = sky_params(sun, 2.5, 128.0)
    ours = frames.sky_cache(p)
    want = oracle.build_sky_cache(empty_view(p), 128)
    assert np.array_equal(ours.view("<u4"), want.view("<u4"))
    c = (np.arange(256, dtype=np.float32) / np.float32(255.0)).astype(np.float32)
    assert np.array_equal(frames.gamma_lut().view("<u4"), oracle.gamma_to_linear(c).view("<u4"))


def test_the_emitter_grid_lists_every_emitter_within_a_cell(oracle):
    # voxel.rs:616-671 on a synthetic world with emitters: from any point, every emitter less
    # than a cell (10 blocks) away along each axis is listed, none two cells or more away.
    spec = frames.WorldSpec(2, 2, 8, cube_types=4, water_level=62, emitter_fraction=0.02)
    scene = frames.build_world(spec)
    assert scene.emitter_count > 5
    p = new_params()
    v = BoundView(p, scene.grid, scene.models, scene.quads, scene.atlas, scene.emitters,
                  scene.emitter_grid, np.zeros(0, "<f4"))
    g = scene.emitter_grid
    cells = int(g[1]) * int(g[2]) * int(g[3])
    lst = g[8 + 2 * cells:]
    pos = scene.emitters["pos"][: scene.emitter_count, :3].astype(int)
    rng = np.random.default_rng(1)
    for _ in range(3000):
        x = rng.integers(-5, 40)
        y = rng.integers(20, 120)
        z = rng.integers(-5, 40)
        start, count = oracle.emitter_cell(v, [x + 0.5, y + 0.5, z + 0.5])
        found = set(lst[start:start + count].tolist())
        dist = np.abs(pos - [x, y, z])
        near = np.all(dist < 10, axis=1)
        far = np.any(dist >= 20, axis=1)
        assert all(i in found for i in np.nonzero(near)[0])
        assert not any(i in found for i in np.nonzero(far)[0])


def test_synthetic_world_layout_decodes_to_the_terrain():
    # The world's brick grid (scene.rs:1-9 layout) holds a solid block exactly where the
    # terrain is, water surfaces at the top water layer, air above.
    spec = frames.WorldSpec(2, 2, 8, cube_types=8, water_level=62, emitter_fraction=0.01,
                            cutout_types=2, cutout_fraction=0.1)
    scene = frames.build_world(spec)
    dims = scene.dims
    off = scene.offsets
    g = scene.grid

    def voxel(x, y, z):
        b = ((y >> 4) * dims[2] + (z >> 4)) * dims[0] + (x >> 4)
        e = int(g[off[0] + b])
        if e == 0xFFFFFFFF:
            return 0
        if e & 0x80000000:
            return e & 0x7FFFFFFF
        local = ((y & 15) * 16 + (z & 15)) * 16 + (x & 15)
        w = int(g[off[1] + e * 2048 + (local >> 1)])
        return (w >> ((local & 1) * 16)) & 0xFFFF

    rng = np.random.default_rng(2)
    water_model = 1 + spec.cube_types + spec.cutout_types
    for _ in range(2000):
        x, z = rng.integers(0, 32, 2)
        h = frames.world_height(spec, int(x), int(z))
        y = int(rng.integers(max(0, h - 20), min(127, h + 5)))
        m = voxel(x, y, z)
        if y <= h:
            # solid: a cube, or an emitter variant, or (buried brick) its uniform stand-in
            assert m != 0 and m != water_model
        elif y < 62:
            assert m == (water_model if y == 61 else 0)
        elif y == h + 1 and m != 0:
            assert spec.cube_types < m <= spec.cube_types + spec.cutout_types
        else:
            assert m == 0


def test_device_exp_is_glibcs_expf(oracle):
    # The device's f32 e^x (water transmittance, math.rs:298-302) follows glibc's expf
    # method; it equals libm's expf (what Rust std calls) on 2^24 inputs spread over every
    # exponent, on the transmittance range, and on the special values.
    import ctypes
    from lumenwright_amd import _native
    rng = np.random.default_rng(7)
    bits = rng.integers(0, 2**32, size=1 << 24, dtype=np.uint64).astype("<u4")
    x = bits.view("<f4")
    x = x[np.isfinite(x)]
    x = np.concatenate([x, -(rng.random(1 << 22) * 104).astype("<f4"),
                        np.array([0.0, -0.0, 88.0, -88.0, 88.72284, 88.72285, -103.97, -103.98,
                                  -104.0, np.inf, -np.inf, 1e-30, -1e-30], "<f4")])
    x = np.ascontiguousarray(x, "<f4")
    out = np.empty_like(x)
    _native.lib().lw_expf_tabulated(ctypes.c_void_p(x.ctypes.data), x.size,
                                    ctypes.c_void_p(out.ctypes.data))
    want = oracle.expf(x)
    differ = out.view("<u4") != want.view("<u4")
    # The two inputs of all 2^32 where the methods part (an exhaustive CPU comparison): the
    # f64 product lies within 1e-15 ulp of an f32 midpoint; one ulp apart.
    known = np.isin(x.view("<u4"), np.array([0x4202422F, 0xC27C65D9], "<u4"))
    assert not (differ & ~known).any(), (x[differ][:5], out[differ][:5], want[differ][:5])
    assert np.all(np.abs(out[known].view("<i4").astype(np.int64)
                         - want[known].view("<i4").astype(np.int64)) <= 1)
    nan_in = np.array([np.nan], "<f4")
    nan_out = np.empty_like(nan_in)
    _native.lib().lw_expf_tabulated(ctypes.c_void_p(nan_in.ctypes.data), 1,
                                    ctypes.c_void_p(nan_out.ctypes.data))
    assert np.isnan(nan_out[0])


def test_map_directions_are_libms_within_an_ulp():
    # map_direction (sky.rs:856-860): (i + 1/2) / w * tau and (j + 1/2) / h * pi - pi/2, their
    # cos and sin in f32 (libm); against f64 of the same f32 angles, within one f32 ulp.
    w, h = 360, 180
    az, el = frames.map_trig(w, h)
    a = ((np.arange(w, dtype="<f4") + np.float32(0.5)) / np.float32(w)) * np.float32(2 * np.pi)
    e = ((np.arange(h, dtype="<f4") + np.float32(0.5)) / np.float32(h)) * np.float32(np.pi) \
        - np.float32(np.pi / 2)
    for got, ang in ((az, a), (el, e)):
        for k, f in ((0, np.cos), (1, np.sin)):
            exact = f(ang.astype(np.float64))
            ulp = np.spacing(np.abs(exact).astype("<f4")).astype(np.float64)
            assert np.all(np.abs(got[:, k] - exact) <= ulp), f


def test_the_blend_is_chunkys_sky_at_weight_one_and_the_map_at_zero(oracle):
    # blend_chunky_sky (sky.rs:836-851): weight 0 leaves every texel, weight 1 makes each
    # the kernel's sky_radiance along the texel's direction (lens map first, i fastest).
    frame = frames.make_frame(frames.config(0, width=8, height=8, spp=1))
    p = frame.params.copy()
    sizes = [40, 20, 36, 18]
    p["sky_map"] = sizes
    s = frame.scene
    v = BoundView(p, s.grid, s.models, s.quads, s.atlas, s.emitters, s.emitter_grid, frame.sky)
    maps = np.random.default_rng(3).random((40 * 20 + 36 * 18) * 3).astype("<f4")
    assert np.array_equal(oracle.blend_chunky_sky(v, 0.0, sizes, maps), maps)
    one = oracle.blend_chunky_sky(v, 1.0, sizes, maps)
    dirs = []
    for w, h in ((40, 20), (36, 18)):
        az, el = frames.map_trig(w, h)
        j, i = np.divmod(np.arange(w * h), w)
        dirs.append(np.stack([el[j, 0] * az[i, 0], el[j, 1], el[j, 0] * az[i, 1]], -1))
    d = np.concatenate(dirs).astype("<f4")
    want = oracle.sky_radiance(v, d)
    assert np.array_equal(one.reshape(-1, 3).view("<u4"), want.reshape(-1, 3).view("<u4"))
    half = oracle.blend_chunky_sky(v, 0.25, sizes, maps)
    lin = np.float32(0.25) * want.reshape(-1) + np.float32(0.75) * maps
    assert np.array_equal(half.view("<u4"), lin.astype("<f4").view("<u4"))
