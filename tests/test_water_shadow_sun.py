"""Closed forms for the oracle's functions that no reference test pins:

- water_interface (crates/lumenwright-kernel/src/integrator.rs:470-504): Schlick's Fresnel
  against the draw it consumes, Snell's law, total internal reflection (which draws nothing),
  the origin offsets and the in-water flag;
- shadow (integrator.rs:316-346): translucent layers as lerp(1, colour, alpha), Beer-Lambert
  transmittance over the segment that leaves the water, opaque blocking, and the eight-hit
  cap;
- sun_texture (integrator.rs:249-277): the sprite texel along a direction, from the angles to
  the sprite's axes, with the v flip, and the windows outside which it is zero.
"""

import numpy as np

from lumenwright_amd.abi import (ALPHA_OPAQUE, ALPHA_TRANSLUCENT, EMITTER, EMITTER_GRID_HEADER, EMPTY,
                                 FLAG_WATER, MODEL, OV_CELL_WORDS, QUAD, f32_bits, new_params)
from scenes import Scene

IOR = 1.333
SIGMA = (0.42, 0.16, 0.10)


def water_params():
    p = new_params()
    p["water_color"] = [*SIGMA, IOR]
    p["misc_f"] = [1.0, 1e-3, 0.0, 0.0]
    return p


def test_water_interface_follows_snell_and_schlick(oracle):
    p = water_params()
    rng = np.random.default_rng(7)
    r0 = ((1 - IOR) / (1 + IOR)) ** 2
    checked = {"reflect": 0, "refract": 0, "tir": 0}
    for k in range(4000):
        front = bool(k % 2)
        n = rng.normal(size=3)
        n /= np.linalg.norm(n)
        cos_i = rng.uniform(0.02, 1.0)
        # d makes angle acos(cos_i) with -n
        t = np.cross(n, rng.normal(size=3))
        t /= np.linalg.norm(t)
        d = -n * cos_i + t * np.sqrt(1 - cos_i ** 2)
        n32, d32 = n.astype("<f4"), d.astype("<f4")
        pos = rng.uniform(0, 10, 3).astype("<f4")
        state = int(rng.integers(0, 2 ** 32))
        d_out, o_out, w, st = oracle.water_interface(p, pos, d32, n32, front, not front, 1e-3, state)
        ci = -float(np.dot(n32.astype(np.float64), d32.astype(np.float64)))
        eta = 1 / IOR if front else IOR
        sin2_t = eta * eta * max(1 - ci * ci, 0.0)
        reflect_dir = d32 + n32 * (2 * ci)
        reflect_dir /= np.linalg.norm(reflect_dir)
        if sin2_t >= 1.0 + 1e-6:
            checked["tir"] += 1
            assert st == state  # no draw
            assert np.allclose(d_out, reflect_dir, atol=1e-5)
            assert np.allclose(o_out, pos + n32 * np.float32(1e-3), atol=1e-6) and w == (not front)
            continue
        if abs(sin2_t - 1.0) < 1e-5:
            continue
        assert st == oracle.pcg_hash(state)  # one draw
        u = float(oracle.rng_draws(state, 1)[0][0])
        fres = r0 + (1 - r0) * (1 - ci) ** 5
        if abs(u - fres) < 1e-5:
            continue
        if u < fres:
            checked["reflect"] += 1
            assert np.allclose(d_out, reflect_dir, atol=1e-5)
            assert np.allclose(o_out, pos + n32 * np.float32(1e-3), atol=1e-6) and w == (not front)
        else:
            checked["refract"] += 1
            # Snell: |d_out x n| = eta |d x n|, across the surface, in the plane of incidence
            assert abs(np.linalg.norm(np.cross(d_out, n32)) - eta * np.sqrt(1 - ci * ci)) < 2e-5
            assert np.dot(d_out, n32) < 0 and abs(np.linalg.norm(d_out) - 1) < 1e-5
            assert abs(np.dot(np.cross(d32, n32), d_out)) < 2e-5
            assert np.allclose(o_out, pos - n32 * np.float32(1e-3), atol=1e-6) and w == front
    assert min(checked.values()) > 100, checked


def test_fresnel_reflects_at_schlicks_rate(oracle):
    # Over many draws the reflected fraction is Schlick's R(cos_i) (r0 = 0.0204 for 1.333).
    p = water_params()
    r0 = ((1 - IOR) / (1 + IOR)) ** 2
    n = np.array([0, 1, 0], "<f4")
    for cos_i in (1.0, 0.5, 0.1):
        d = np.array([np.sqrt(1 - cos_i ** 2), -cos_i, 0], "<f4")
        reflected = sum(oracle.water_interface(p, np.zeros(3), d, n, True, False, 1e-3, s)[0][1] > 0
                        for s in range(6000))
        want = r0 + (1 - r0) * (1 - cos_i) ** 5
        assert abs(reflected / 6000 - want) < 4 * np.sqrt(want * (1 - want) / 6000) + 1e-3, (cos_i, reflected)


def row_scene(layers, atlas):
    """One 16^3 brick with no blocks; overlay quads facing +x or -x at x = k + 0.5 across the
    row of cells (k, 5, 5), y and z in [5, 6]. layers: (k, flags, texel index, facing +x)."""
    quads, cells = [], {}
    for k, flags, tex, plus in layers:
        q = np.zeros((), dtype=QUAD)
        if plus:  # normal = ev x eu = +x
            q["p0"] = [k + 0.5, 5.0, 5.0, f32_bits(flags)]
            q["eu"] = [0.0, 0.0, 1.0, f32_bits(0xFFFFFFFF)]
            q["ev"] = [0.0, 1.0, 0.0, 0.0]
        else:
            q["p0"] = [k + 0.5, 5.0, 5.0, f32_bits(flags)]
            q["eu"] = [0.0, 1.0, 0.0, f32_bits(0xFFFFFFFF)]
            q["ev"] = [0.0, 0.0, 1.0, 0.0]
        q["uv"] = [tex + 0.5, 0.5, 0.0, 0.0]
        q["uvt"] = [0.0, 0.0, 0.0, 0.0]
        cells.setdefault(k, []).append(len(quads))
        quads.append(q)
    ov_cells = np.zeros(OV_CELL_WORDS, dtype="<u4")
    ov_quads = []
    for k, qs in sorted(cells.items()):
        cell = (5 * 16 + 5) * 16 + k
        ov_cells[cell * 2] = len(ov_quads)
        ov_cells[cell * 2 + 1] = len(qs)
        ov_quads += qs
    grid = []
    off_top = len(grid); grid += [EMPTY]
    off_bricks = len(grid)
    off_ov_top = len(grid); grid += [0]
    off_ov_cells = len(grid); grid += list(ov_cells)
    off_ov_quads = len(grid); grid += ov_quads
    grid.append(0)
    p = water_params()
    p["grid_dims"] = [1, 1, 1, 1]
    p["off"] = [off_top, off_bricks, off_ov_top, off_ov_cells]
    p["off2"] = [off_ov_quads, len(atlas), 1, len(quads)]
    return Scene(p, grid, np.zeros(1, MODEL), quads, atlas, np.zeros(1, EMITTER),
                 np.zeros(EMITTER_GRID_HEADER, "<u4"), [])


ATLAS = np.array([0x800000FF,               # 0: red (255, 0, 0), alpha 128
                  0xFF00FF00,                # 1: opaque green
                  0x40FFFFFF], "<u4")        # 2: white, alpha 64


def run_shadow(oracle, layers, tmax=100.0):
    v = row_scene(layers, ATLAS).view()
    return oracle.shadow(v, [[0.25, 5.5, 5.5]], [[1.0, 0.0, 0.0]], tmax)[0].astype(np.float64)


def test_shadow_through_translucent_layers_and_water(oracle):
    a, a2 = 128 / 255, 64 / 255
    eps = 1e-3
    red = (1 - 0, 1 - a, 1 - a)  # lerp(1, (1, 0, 0), a)
    assert np.allclose(run_shadow(oracle, []), 1.0)
    assert np.allclose(run_shadow(oracle, [(2, ALPHA_TRANSLUCENT, 0, True)]), red, rtol=1e-6)
    two = run_shadow(oracle, [(2, ALPHA_TRANSLUCENT, 0, True), (4, ALPHA_TRANSLUCENT, 2, False)])
    assert np.allclose(two, np.array(red) * (1 - a2 + a2 * 1.0), rtol=1e-6)  # white at alpha 64: 1
    assert np.allclose(run_shadow(oracle, [(6, ALPHA_OPAQUE, 1, True)]), 0.0)
    # Leaving the water (a water quad whose normal points along the ray) absorbs over the
    # segment from the last origin (the previous hit plus eps) to the surface; entering it
    # (normal against the ray) absorbs nothing.
    leave = run_shadow(oracle, [(3, ALPHA_TRANSLUCENT, 0, True), (9, FLAG_WATER, 1, True)])
    seg = 9.5 - (3.5 + eps)
    assert np.allclose(leave, np.array(red) * np.exp(-np.array(SIGMA) * seg), rtol=2e-6)
    enter = run_shadow(oracle, [(9, FLAG_WATER, 1, False)])
    assert np.allclose(enter, 1.0)
    first = run_shadow(oracle, [(9, FLAG_WATER, 1, True)])
    assert np.allclose(first, np.exp(-np.array(SIGMA) * (9.5 - 0.25)), rtol=2e-6)
    # tmax short of a layer: not counted
    assert np.allclose(run_shadow(oracle, [(2, ALPHA_TRANSLUCENT, 0, True)], tmax=2.0), 1.0)


def test_shadow_stops_after_eight_hits(oracle):
    a = 128 / 255
    eight = [(k, ALPHA_TRANSLUCENT, 0, True) for k in range(1, 9)]
    want8 = np.array([1.0, (1 - a) ** 8, (1 - a) ** 8])
    assert np.allclose(run_shadow(oracle, eight), want8, rtol=1e-5)
    # a ninth layer, translucent or opaque, is never traced
    assert np.allclose(run_shadow(oracle, eight + [(10, ALPHA_TRANSLUCENT, 0, True)]), want8, rtol=1e-5)
    assert np.allclose(run_shadow(oracle, eight + [(10, ALPHA_OPAQUE, 1, True)]), want8, rtol=1e-5)
    assert np.allclose(run_shadow(oracle, eight[:7] + [(10, ALPHA_OPAQUE, 1, True)]), 0.0)


def sun_view(brightness=2.0, width=0.05, sprite=(3, 1, 4, 4), sun_dir=(0.3, 0.8, 0.52)):
    """A 10 x 6 atlas whose texel (x, y) has red 10 + 12 (y * 10 + x); the sun sprite the
    4 x 4 block at (3, 1); the sun along sun_dir with axes su, sv perpendicular to it."""
    p = new_params()
    s = np.array(sun_dir, np.float64)
    s /= np.linalg.norm(s)
    su = np.cross(s, [0, 0, 1.0])
    su /= np.linalg.norm(su)
    sv = np.cross(su, s)
    p["sun_dir"] = [*s.astype("<f4"), 0.0]
    p["sun_su"] = [*su.astype("<f4"), brightness]
    p["sun_sv"] = [*sv.astype("<f4"), width]
    p["sun_sprite"] = list(sprite)
    idx = np.arange(60)
    atlas = ((10 + 3 * idx) | (255 << 24)).astype("<u4")
    p["off2"] = [0, 10, 6, 0]
    sc = Scene(p, [0], np.zeros(1, MODEL), np.zeros(1, QUAD), atlas, np.zeros(1, EMITTER),
               np.zeros(EMITTER_GRID_HEADER, "<u4"), [])
    return sc, s, su, sv


def test_sun_texture_reads_the_sprite_texel(oracle):
    width, brightness = 0.05, 2.0
    sc, s, su, sv = sun_view(brightness, width)
    v = sc.view()
    lin = oracle.gamma_to_linear(np.arange(10, 190, 3, dtype="<f4") / np.float32(255.0)).astype(np.float64)
    rng = np.random.default_rng(3)
    checked = 0
    for _ in range(2000):
        # angles a, b in [0, 2 width): the direction makes angle a - width with the plane
        # perpendicular to su (asin(d . su) = a - width), and likewise b with sv
        a, b = rng.uniform(-0.2 * width, 2.2 * width, 2)
        du, dv = np.sin(a - width), np.sin(b - width)
        d = s * np.sqrt(max(1 - du * du - dv * dv, 0)) + su * du + sv * dv
        got = oracle.sun_texture(v, d.astype("<f4")[None])[0].astype(np.float64)
        inside = 0 <= a < 2 * width and 0 <= b < 2 * width
        fx, fy = a / (2 * width) * 4, (1 - b / (2 * width)) * 4
        near_edge = min(abs(fx - round(fx)), abs(fy - round(fy))) < 1e-3 or min(
            abs(a), abs(a - 2 * width), abs(b), abs(b - 2 * width)) < 1e-5
        if near_edge:
            continue
        checked += 1
        if not inside:
            assert np.all(got == 0.0), (a, b, got)
            continue
        x, y = int(np.floor(fx)), int(np.floor(fy))
        want = lin[(1 + y) * 10 + 3 + x] * brightness
        assert abs(got[0] - want) < 1e-6 * max(1.0, want) and got[1] == 0.0, (a, b, x, y, got)
    assert checked > 1500
    # facing away (d . sun < 0.5) and a zero-width sprite: no texture
    assert np.all(oracle.sun_texture(v, (-s).astype("<f4")[None]) == 0.0)
    sc0, *_ = sun_view(sprite=(3, 1, 0, 4))
    assert np.all(oracle.sun_texture(sc0.view(), s.astype("<f4")[None]) == 0.0)


def water_plane_scene(cos_i):
    """The reference's map test scene (tests.rs:804-831: lens map 5.0, light map 0.3) with
    Chunky's water plane at y = 10 drawn, and a 1-pixel camera outside the grid's columns
    (x < 0) looking down onto the plane at incidence acos(cos_i)."""
    import scenes as sc
    s = sc.test_scene(False)
    p = s.params
    p["sky_map"] = [4.0, 2.0, 2.0, 2.0]
    s.sky = np.concatenate([np.full(24, 5.0), np.full(12, 0.3)]).astype("<f4")
    p["water_plane"] = [10.0, 1.0, 0.0, 0.0]
    p["water_color"] = [*SIGMA, IOR]
    p["frame"] = [1, 1, 0, 1]
    sin_i = np.sqrt(1 - cos_i ** 2)
    fwd = np.array([-sin_i, -cos_i, 0.0], "<f4")
    right = np.array([0.0, 0.0, 1.0], "<f4")
    up = np.cross(right, fwd).astype("<f4")
    p["cam_pos"] = [-30.0, 40.0, 8.0, 1e-4]  # tan(fov / 2): a near-pinhole 1-pixel frame
    p["cam_fwd"] = [*fwd, 1.0]
    p["cam_right"] = [*right, 0.0]
    p["cam_up"] = [*up, 0.0]
    return s


def test_the_water_plane_reflects_at_schlicks_rate(oracle):
    # integrator.rs:506-534 and :577-590. A camera ray meets the plane outside the grid's
    # columns. Reflected, it leaves upward and reads the lens map (5.0). Refracted, it stays
    # under the plane and in the water, and no sky is added. The mean over the samples is
    # R(cos_i) * 5.0.
    r0 = ((1 - IOR) / (1 + IOR)) ** 2
    n = 4096
    for cos_i in (0.9, 0.4, 0.15):
        s = water_plane_scene(cos_i)
        c = oracle.render_samples(s.view(), np.zeros(n), np.zeros(n), np.arange(n)).astype(np.float64)
        assert np.all((np.abs(c - 5.0) < 1e-5) | (c == 0.0))  # reflected: the lens map; refracted: nothing
        frac = (c[:, 0] > 0).mean()
        want = r0 + (1 - r0) * (1 - cos_i) ** 5
        assert abs(frac - want) < 4 * np.sqrt(want * (1 - want) / n) + 2e-3, (cos_i, frac, want)
    # not drawn (water_plane.y = 0): the ray passes the plane and reads the lens map
    s = water_plane_scene(0.4)
    s.params["water_plane"] = [10.0, 0.0, 0.0, 0.0]
    c = oracle.render_samples(s.view(), np.zeros(16), np.zeros(16), np.arange(16))
    assert np.all(np.abs(c - 5.0) < 1e-5)
