"""Quick mode: the oracle's restatement of the reference's rasterizer (crates/lumenwright/src/
render/cpu/raster.rs) and atlas sampling (crates/lumenwright/src/render/texture.rs), pinned by
the reference's own texture test (texture.rs:147-177), closed forms of the sampler (mip level
selection, trilinear weights, the anisotropic taps) and of the rasterizer (coverage against
exact point-in-triangle tests, back-face culling, depth order, the alpha test, blending,
see-through drawing, near-plane clipping, both fogs, the sky behind the geometry), and libm's
log2f band that the device's mip level relies on."""

import numpy as np
import pytest

from lumenwright_amd import frames
from lumenwright_amd.abi import QUICK_VERTEX, QuickFrame, QuickMesh

NO_FOG = (1.0e6, 2.0e6)


# ---- the sampler (texture.rs) ----

def test_magnified_texels_stay_sharp(oracle):
    # texture.rs:151-176: a 2-level 4 x 4 atlas with a checker; 100 pixels per texel read the
    # texel exactly, 4 texels per pixel read level 2 (the mip colour).
    level0 = np.array([0xFFFFFFFF if (i % 4 + i // 4) % 2 == 0 else 0xFF000000 for i in range(16)], "<u4")
    mips = np.full(4 + 1, 0xFF808080, "<u4")
    d, e = (0.01 / 4.0, 0.0), (0.0, 0.01 / 4.0)

    def block(uv, du, dv):
        return oracle.atlas_sample(level0, mips, (4, 4), 2, 4, [[*uv, *du, *dv]], mode=1)[0]

    assert abs(block((0.3 / 4, 0.3 / 4), d, e)[0] - 1.0) < 1e-5
    assert abs(block((1.3 / 4, 0.3 / 4), d, e)[0]) < 1e-5
    c = block((0.5, 0.5), (d[0] * 400, 0.0), (0.0, e[1] * 400))
    assert abs(c[0] - 128.0 / 255.0) < 1e-5


def levels_atlas(w=16, h=16, levels=3, grey=(40, 120, 200, 250)):
    """Every texel of level l the grey grey[l]."""
    def g(v):
        return np.uint32(v | (v << 8) | (v << 16) | (255 << 24))
    level0 = np.full(w * h, g(grey[0]), "<u4")
    mips = np.concatenate([np.full((w >> l) * (h >> l), g(grey[l]), "<u4") for l in range(1, levels + 1)])
    return level0, mips


@pytest.mark.parametrize("s,lod", [(1.0, 0.0), (2.0, 1.0), (4.0, 2.0), (2 ** 0.5, 0.5), (2 ** 1.25, 1.25),
                                   (64.0, 3.0), (0.25, 0.0)])
def test_the_footprint_picks_the_mip_level(oracle, s, lod):
    # An isotropic footprint of s texels: one tap at level log2(s) clamped to [0, levels],
    # trilinear between the levels' greys.
    w = h = 16
    level0, mips = levels_atlas(w, h, 3)
    q = [[0.5, 0.5, s / w, 0.0, 0.0, s / h]]
    got = oracle.atlas_sample(level0, mips, (w, h), 3, h, q, mode=2)[0]
    greys = np.array([40, 120, 200, 250]) / 255.0
    lo = int(np.floor(lod))
    want = greys[lo] if lo == lod else greys[lo] * (1 - (lod - lo)) + greys[lo + 1] * (lod - lo)
    assert abs(got[0] - want) < 1e-5, (got, want)
    assert abs(got[3] - 1.0) < 1e-6


@pytest.mark.parametrize("ratio,taps,lod", [(4, 4, 0.0), (2, 2, 0.0), (8, 4, 1.0), (3, 3, 0.0)])
def test_anisotropic_taps_along_the_major_axis(oracle, ratio, taps, lod):
    # Column k of level 0 has grey 10 k; a footprint of `ratio` texels along u and one along v:
    # N = min(ceil(ratio), 4) taps spaced evenly over the u axis at level log2(ratio / N). At a
    # texel centre the taps fall at offsets (i + 0.5) / N - 0.5 of the footprint.
    w, h = 32, 8
    cols = np.arange(w) * 7
    level0 = np.tile((cols | (cols << 8) | (cols << 16) | (255 << 24)).astype("<u4"), h)
    mips = np.concatenate([np.full((w >> l) * (h >> l), 0xFF000000, "<u4") for l in (1, 2, 3)])
    k = 12
    u = (k + 0.5) / w
    got = oracle.atlas_sample(level0, mips, (w, h), 3, h, [[u, 0.5, ratio / w, 0.0, 0.0, 1.0 / h]], mode=2)[0]
    want = 0.0
    for i in range(taps):
        t = (k + 0.5) + ratio * ((i + 0.5) / taps - 0.5) - 0.5  # bilinear texel coordinate
        f = t - np.floor(t)
        a, b = cols[int(np.floor(t))], cols[int(np.floor(t)) + 1]
        want += (a * (1 - f) + b * f) / 255.0
    want /= taps
    if lod > 0:
        want *= (1 - (lod - np.floor(lod))) if lod % 1 else 0.0  # level 1 is black
    assert abs(got[0] - want) < 2e-6, (got, want)


def test_sample_block_snaps_to_texel_centres(oracle):
    # One texel per pixel or more: sampleNearest moves the coordinate to the texel centre
    # unless within half a screen pixel of a texel edge.
    w = h = 16
    vals = (np.arange(w * h) * 37) % 256
    level0 = (vals | (vals << 8) | (vals << 16) | (255 << 24)).astype("<u4")
    _, mips = levels_atlas(w, h, 2)
    rng = np.random.default_rng(4)
    tx, ty = rng.integers(0, w, 200), rng.integers(0, h, 200)
    fx, fy = rng.uniform(0.3, 0.7, 200), rng.uniform(0.3, 0.7, 200)
    scale = 0.25  # texels per pixel: a quarter pixel from an edge is still inside
    q = np.stack([(tx + fx) / w, (ty + fy) / h, np.full(200, scale / w), np.zeros(200), np.zeros(200),
                  np.full(200, scale / h)], -1)
    got = oracle.atlas_sample(level0, mips, (w, h), 2, h, q, mode=0)
    assert np.array_equal(np.rint(got[:, 0] * 255).astype(int), vals[ty * w + tx])


def test_nearest_outside_the_mipmapped_rows(oracle):
    w, h = 8, 8
    vals = np.arange(w * h)
    level0 = (vals | (255 << 24)).astype("<u4")
    _, mips = levels_atlas(w, h, 1)
    rng = np.random.default_rng(1)
    uv = rng.uniform(-0.2, 1.2, (300, 2))
    q = np.concatenate([uv, rng.uniform(0, 1, (300, 4))], 1)
    got = oracle.atlas_sample(level0, mips, (w, h), 1, 4, q, mode=0)  # rows 0..4 mipmapped
    f32 = uv.astype("<f4")
    x = np.clip(np.floor(f32[:, 0] * np.float32(w)), 0, w - 1).astype(int)
    y = np.clip(np.floor(f32[:, 1] * np.float32(h)), 0, h - 1).astype(int)
    lower = f32[:, 1] * np.float32(h) >= 4
    assert lower.sum() > 100
    assert np.array_equal(np.rint(got[lower, 0] * 255).astype(int), vals[y * w + x][lower])


def test_log2f_band_holds_for_the_mip_level(oracle):
    # The device's log2 (log2_libm): every f32 in [1, 65536] where libm's log2f differs from
    # the rounded f64 log2 lies within the absolute band 2^-28 of a rounding midpoint.
    dev, outside, band, not_nb, max_dist = oracle.log2f_band()
    assert dev > 0 and outside == 0 and not_nb == 0
    assert max_dist < 0.75 * 2.0 ** -28  # 2^-29.0
    assert band < 1 << 24


# ---- the rasterizer (raster.rs) ----

def plain_sky(w, h, fog_colour=(0.5, 0.5, 0.5), render_fog=NO_FOG, env_fog=NO_FOG, yaw=180.0, pitch=0.0,
              fov=90.0):
    """A sky of one colour (no disc, no sunrise, no bodies, no stars) seen by a camera at the
    origin looking down -z (yaw 180)."""
    d = frames.sky_description(sky=(0.1, 0.2, 0.3), fog=fog_colour, sky_fog_end=1.0e6, render_fog=render_fog,
                               env_fog=env_fog, sky_disc=False, sunrise=(0, 0, 0, 0), sun_angle=0.0)
    return frames.quick_sky_new(frames.camera_rays(yaw, pitch, fov, w, h), w, h, d)


def small_atlas():
    """8 x 8 texels, not mipmapped: the colours a pixel of a triangle may take."""
    px = np.zeros((8, 8), "<u4")
    px[0:4, 0:4] = 0xFF3264C8      # (200, 100, 50, 255)
    px[0:4, 4:8] = 0xFF00FF00      # green
    px[4:8, 0:4] = 0x64FFFFFF      # white, alpha 100 (cut by the alpha test)
    px[4:8, 4:8] = 0x800000FF      # red, alpha 128
    return px.reshape(-1)


RECTS = {"solid": (0, 0), "green": (4, 0), "cut": (0, 4), "red": (4, 4)}


def quad(p0, e1, e2, rect, colour=(255, 255, 255, 255), size=8):
    """A quad p0, p0 + e1, p0 + e1 + e2, p0 + e2 (counter-clockwise seen from the side e1 x e2
    points to) on a 4 x 4 texel rect of the test atlas."""
    p0, e1, e2 = (np.asarray(a, np.float64) for a in (p0, e1, e2))
    x, y = RECTS[rect]
    v = np.zeros(4, QUICK_VERTEX)
    v["pos"] = np.stack([p0, p0 + e1, p0 + e1 + e2, p0 + e2]).astype("<f4")
    u0, v0, u1, v1 = (x + 0.5) / size, (y + 0.5) / size, (x + 3.5) / size, (y + 3.5) / size
    v["uv"] = [[u0, v1], [u1, v1], [u1, v0], [u0, v0]]
    v["color"] = colour
    return v


def facing(z, x0=-50.0, x1=50.0, y0=-50.0, y1=50.0, **kw):
    """A quad in the plane z facing the camera at the origin (normal +z)."""
    return quad((x0, y0, z), (x1 - x0, 0, 0), (0, y1 - y0, 0), **kw)


def mesh_of(opaque=(), translucent=(), see_through=()):
    verts, lists, n = [], [], 0
    for group in (opaque, translucent, see_through):
        idx = []
        for q in group:
            verts.append(q)
            idx += [n, n + 1, n + 2, n, n + 2, n + 3]
            n += 4
        lists.append(np.array(idx, "<u4"))
    v = np.concatenate(verts) if verts else np.zeros(0, QUICK_VERTEX)
    return QuickMesh(vertices=v, opaque=lists[0], translucent=lists[1], see_through=lists[2])


def frame_of(mesh, w=64, h=48, sky=None, atlas=None, atlas_size=(8, 8), mips=None, levels=0, mip_rows=0,
             fov=90.0, near=0.05, far=1000.0):
    return QuickFrame(mesh=mesh, atlas_pixels=small_atlas() if atlas is None else atlas, atlas_size=atlas_size,
                      atlas_mips=np.zeros(0, "<u4") if mips is None else mips, mip_levels=levels,
                      mip_rows=mip_rows, stars=frames.star_table(), width=w, height=h,
                      view_proj=frames.quick_view_proj(180.0, 0.0, fov, w / h, near, far),
                      sky=sky if sky is not None else plain_sky(w, h, fov=fov))


def channels(img, w, h):
    img = np.asarray(img, "<u4").reshape(h, w)
    return np.stack([(img >> s) & 0xFF for s in (0, 8, 16, 24)], -1).astype(int)


def pixel_rays(w, h, fov=90.0):
    """The camera-space direction through each pixel centre (the camera looks down -z)."""
    t = np.tan(np.radians(fov) / 2)
    x = ((np.arange(w) + 0.5) / w * 2 - 1) * t * (w / h)
    y = (1 - (np.arange(h) + 0.5) / h * 2) * t
    X, Y = np.meshgrid(x, y)
    return np.stack([X, Y, -np.ones_like(X)], -1)


def test_a_solid_quad_is_texel_times_vertex_colour(oracle):
    w, h = 64, 48
    img = channels(oracle.rasterize(frame_of(mesh_of([facing(-5.0, rect="solid", colour=(255, 128, 64, 255))]))), w, h)
    # (200, 100, 50) / 255 times (255, 128, 64) / 255, times 255, rounded: 200, 50.2, 12.5
    assert np.all(img[..., 0] == 200) and np.all(img[..., 1] == 50) and np.all(img[..., 2] == 13)
    assert np.all(img[..., 3] == 255)


def test_an_empty_mesh_is_the_sky(oracle):
    # Each pixel is QuickSky::background along pixel_dir (sky.rs:380-388), packed (raster.rs:
    # 265-278): the dusk sky with its disc, sunrise fan, sun and stars.
    from test_game_sky import dusk_description
    w, h = 48, 27
    cam = frames.camera_rays(-120.0, -10.0, 70.0, w, h)
    q = frames.quick_sky_new(cam, w, h, dusk_description(True))
    atlas = np.full(16, 0xFFFFFFFF, "<u4")
    f = QuickFrame(mesh=mesh_of(), atlas_pixels=atlas, atlas_size=(4, 4), atlas_mips=np.zeros(0, "<u4"),
                   mip_levels=0, mip_rows=0, stars=frames.star_table(), width=w, height=h,
                   view_proj=frames.quick_view_proj(-120.0, -10.0, 70.0, w / h, 0.05, 1000.0), sky=q)
    got = oracle.rasterize(f)
    m = np.array(q.cam_misc, "<f4")
    fw, r, u = (np.array(a, "<f4")[:3] for a in (q.cam_forward, q.cam_right, q.cam_up))
    X, Y = np.meshgrid(np.arange(w, dtype="<f4"), np.arange(h, dtype="<f4"))
    two, one, half = np.float32(2), np.float32(1), np.float32(0.5)
    sx = ((two * (X + half) / m[2] - one) * m[1]) * m[0]
    sy = (one - two * (Y + half) / m[3]) * m[0]
    d = fw + r * sx[..., None] + u * sy[..., None]
    length = np.sqrt((d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]) + d[..., 2] * d[..., 2])
    d = d * (one / length)[..., None]
    bg = oracle.quick_background(q, d.reshape(-1, 3), atlas, 4)
    c = np.floor(np.clip(bg, 0, 1) * np.float32(255) + np.float32(0.5)).astype(np.uint32)
    want = c[:, 0] | (c[:, 1] << 8) | (c[:, 2] << 16) | np.uint32(0xFF000000)
    assert np.array_equal(got, want)
    assert len(np.unique(got)) > 20


def test_coverage_is_the_pixel_centres_inside(oracle):
    # Random front-facing triangles against an exact point-in-triangle test of the pixel
    # centres (pixels within 1e-3 px of an edge are not checked); ragged 75 x 41 frame.
    w, h = 75, 41
    rng = np.random.default_rng(11)
    rays = pixel_rays(w, h)
    checked = 0
    for _ in range(12):
        z = -rng.uniform(2, 20)
        pts = np.concatenate([rng.uniform(-1.2, 1.2, (3, 2)) * -z, np.full((3, 1), z)], 1)
        # screen positions in f64 (y down), counter-clockwise in camera space
        e = np.cross(pts[1] - pts[0], pts[2] - pts[0])
        if e[2] < 0:
            pts = pts[[0, 2, 1]]
        v = np.zeros(3, QUICK_VERTEX)
        v["pos"] = pts.astype("<f4")
        v["uv"] = (0.1, 0.1)
        v["color"] = (255, 255, 255, 255)
        m = QuickMesh(vertices=v, opaque=np.array([0, 1, 2], "<u4"))
        img = channels(oracle.rasterize(frame_of(m, w, h)), w, h)
        drawn = img[..., 0] == 200
        t = np.tan(np.radians(45.0))
        sx = (pts[:, 0] / -pts[:, 2] / (t * w / h) * 0.5 + 0.5) * w
        sy = (0.5 - pts[:, 1] / -pts[:, 2] / t * 0.5) * h
        X, Y = np.meshgrid(np.arange(w) + 0.5, np.arange(h) + 0.5)
        def edge(a, b):
            return (sx[b] - sx[a]) * (Y - sy[a]) - (sy[b] - sy[a]) * (X - sx[a])
        # counter-clockwise in camera space is clockwise on the y-down screen: all edges <= 0
        es = np.stack([edge(0, 1), edge(1, 2), edge(2, 0)])
        lens = np.array([np.hypot(sx[b] - sx[a], sy[b] - sy[a]) for a, b in ((0, 1), (1, 2), (2, 0))])
        dist = es / lens[:, None, None]
        inside = np.all(dist <= 0, axis=0)
        clear = np.all(np.abs(dist) > 1e-3, axis=0)
        assert np.array_equal(drawn[clear], inside[clear])
        checked += int(inside[clear].sum())
    assert checked > 500


def test_back_faces_are_culled(oracle):
    front = facing(-5.0, rect="solid")
    back = front[[0, 3, 2, 1]]  # the same quad wound the other way
    img = channels(oracle.rasterize(frame_of(mesh_of([back]))), 64, 48)
    assert np.all(img[..., :3] == 128)


@pytest.mark.parametrize("order", ["near first", "far first"])
def test_the_nearer_opaque_surface_wins(oracle, order):
    near = facing(-4.0, x0=-10, x1=1, rect="solid")
    far = facing(-6.0, x0=-1, x1=10, rect="green")
    qs = [near, far] if order == "near first" else [far, near]
    img = channels(oracle.rasterize(frame_of(mesh_of(qs))), 64, 48)
    # left of the screen centre: the near quad; where only the far quad is: green
    assert np.all(img[:, :28, 0] == 200)
    assert np.all(img[:, 40:, 1] == 255) and np.all(img[:, 40:, 0] == 0)
    overlap = img[:, 32:34]  # both quads: the near one
    assert np.all(overlap[..., 0] == 200)


def test_the_alpha_test_cuts_below_one_half(oracle):
    img = channels(oracle.rasterize(frame_of(mesh_of([facing(-5.0, rect="cut")]))), 64, 48)
    assert np.all(img[..., :3] == 128)  # alpha 100 / 255 < 0.5: the sky
    img = channels(oracle.rasterize(frame_of(mesh_of([facing(-5.0, rect="red")]))), 64, 48)
    assert np.all(img[..., 0] == 255) and np.all(img[..., 1] == 0)  # alpha 128 / 255 >= 0.5


def blend(dst, src, a):
    return dst * (1 - a) + src * a


def test_translucent_surfaces_blend_over_what_is_behind(oracle):
    a = 128 / 255  # red texel alpha, vertex alpha 255
    # over the sky (0.5 grey)
    img = channels(oracle.rasterize(frame_of(mesh_of(translucent=[facing(-5.0, rect="red")]))), 64, 48)
    want = np.floor(np.array([blend(0.5 * 255 / 255, 1.0, a), blend(128 / 255, 0.0, a)]) * 255 + 0.5)
    assert np.all(img[..., 0] == want[0]) and np.all(img[..., 1] == want[1])
    # vertex alpha scales it: 0.25 * 128 / 255
    img = channels(oracle.rasterize(frame_of(mesh_of(
        translucent=[facing(-5.0, rect="red", colour=(255, 255, 255, 64))]))), 64, 48)
    a2 = (128 / 255) * (64 / 255)
    assert np.all(np.abs(img[..., 0] - np.floor(blend(128 / 255, 1.0, a2) * 255 + 0.5)) <= 1)
    # two layers, back to front, over an opaque green wall; a translucent quad behind the wall
    # is hidden (depth test) and writes no depth itself
    wall = facing(-8.0, rect="green")
    hidden = facing(-9.0, rect="red")
    l1 = facing(-6.0, rect="red")
    l2 = facing(-5.0, rect="red", colour=(255, 255, 255, 64))
    img = channels(oracle.rasterize(frame_of(mesh_of([wall], translucent=[hidden, l1, l2]))), 64, 48)
    def q8(x):  # each layer's result is packed to 8 bits before the next blends over it
        return np.floor(x * 255 + 0.5) / 255
    g = blend(q8(blend(1.0, 0.0, a)), 0.0, a2)
    r = blend(q8(blend(0.0, 1.0, a)), 1.0, a2)
    # The quads' shared diagonal runs through pixel centres here (one a row): where a rounded
    # edge value is exactly zero both triangles cover the pixel (the edge test keeps zero) and
    # blend the layer twice, so the diagonal is not checked here.
    Y, X = np.mgrid[0:48, 0:64]
    off = np.abs(X - (55 - Y)) > 1
    assert np.all(np.abs(img[off, 1] - g * 255) <= 1) and np.all(np.abs(img[off, 0] - r * 255) <= 1)
    assert np.any(img[~off, 1] < g * 255 - 1)  # blended twice


def test_see_through_is_drawn_over_everything(oracle):
    a = 128 / 255
    wall = facing(-3.0, rect="green")
    text = facing(-9.0, rect="red")  # behind the wall, drawn over it anyway
    img = channels(oracle.rasterize(frame_of(mesh_of([wall], see_through=[text]))), 64, 48)
    Y, X = np.mgrid[0:48, 0:64]
    off = np.abs(X - (55 - Y)) > 1  # not the shared diagonal (see the translucent test)
    assert np.all(np.abs(img[off, 0] - blend(0.0, 1.0, a) * 255) <= 0.5 + 1e-9)
    assert np.all(np.abs(img[off, 1] - blend(1.0, 0.0, a) * 255) <= 0.5 + 1e-9)
    # the same quad in the translucent pass is hidden behind the wall
    img = channels(oracle.rasterize(frame_of(mesh_of([wall], translucent=[text]))), 64, 48)
    assert np.all(img[..., 1] == 255) and np.all(img[..., 0] == 0)


def test_near_plane_clipping_keeps_what_is_in_front(oracle):
    # A floor at y = -1 from behind the camera (z = +5) to z = -50: one triangle has one vertex
    # in front of the near plane (clipped to a triangle), the other two (clipped to a quad).
    w, h = 64, 48
    floor = quad((-30.0, -1.0, 5.0), (60.0, 0, 0), (0, 0, -55.0), rect="solid")  # facing +y
    img = channels(oracle.rasterize(frame_of(mesh_of([floor]), w, h)), w, h)
    rays = pixel_rays(w, h)
    with np.errstate(divide="ignore", invalid="ignore"):
        t = -1.0 / rays[..., 1]
    hit = (rays[..., 1] < 0) & (t * -rays[..., 2] < 50.0) & (np.abs(t * rays[..., 0]) < 30.0)
    near_edge = np.abs(t * -rays[..., 2] - 50.0) < 0.5
    drawn = img[..., 0] == 200
    assert np.array_equal(drawn[~near_edge], hit[~near_edge])
    assert drawn[-1].all()  # the bottom row: the floor just in front of the camera


@pytest.mark.parametrize("which", ["render", "environment"])
def test_fog_closes_over_distance(oracle, which):
    # Render fog on the cylindrical distance max(sqrt(x^2 + z^2), |y|), environment fog on the
    # spherical one; the larger factor wins (sky.rs:536-544).
    w, h = 64, 48
    fog = (0.2, 0.9, 0.4)
    start, end = 4.0, 14.0
    sky = plain_sky(w, h, fog_colour=fog, render_fog=(start, end) if which == "render" else NO_FOG,
                    env_fog=(start, end) if which == "environment" else NO_FOG)
    img = channels(oracle.rasterize(frame_of(mesh_of([facing(-6.0, rect="solid")]), w, h, sky=sky)), w, h)
    rays = pixel_rays(w, h)
    pos = rays * 6.0
    if which == "render":
        dist = np.maximum(np.hypot(pos[..., 0], pos[..., 2]), np.abs(pos[..., 1]))
    else:
        dist = np.linalg.norm(pos, axis=-1)
    f = np.clip((dist - start) / (end - start), 0, 1)
    src = np.array([200, 100, 50]) / 255
    want = (src * (1 - f[..., None]) + np.array(fog) * f[..., None]) * 255
    assert np.all(np.abs(img[..., :3] - want) <= 1.0)
    assert np.ptp(f) > 0.1


@pytest.mark.parametrize("level", [0, 1, 2])
def test_the_rasterizer_reads_the_mip_level_of_its_footprint(oracle, level):
    # A facing 32 x 32 pixel quad on a 256 x 256 atlas (every level one grey) whose UVs span
    # 3 * 2^level texels per pixel across and 2.5 times fewer down: the pixel derivatives
    # (raster.rs:316-323) give N = ceil(2.5) = 3 taps at level log2(3 * 2^level / 3)
    # (texture.rs:102-119). Outside the mipmapped rows (mip_rows 0) the quad reads level 0.
    w = h = 64
    greys = (40, 90, 140, 190, 240)
    level0, mips = levels_atlas(256, 256, 4, greys)
    pmax = 3.0 * 2 ** level
    U, V = pmax * 32 / 256, pmax / 2.5 * 32 / 256
    v = quad((-5.0, -5.0, -10.0), (10.0, 0, 0), (0, 10.0, 0), rect="solid")  # 32 x 32 pixels
    v0 = 0.05
    v["uv"] = [[0, v0 + V], [U, v0 + V], [U, v0], [0, v0]]
    m = mesh_of([v])
    img = channels(oracle.rasterize(frame_of(m, w, h, atlas=level0, atlas_size=(256, 256), mips=mips, levels=4,
                                             mip_rows=256)), w, h)
    inner = img[18:46, 18:46]
    assert np.all(np.abs(inner[..., 0] - greys[level]) <= 1), inner[..., 0]
    img = channels(oracle.rasterize(frame_of(m, w, h, atlas=level0, atlas_size=(256, 256), mips=mips, levels=4,
                                             mip_rows=0)), w, h)
    assert np.all(img[18:46, 18:46, 0] == greys[0])


def test_a_synthetic_world_frame_is_drawn(oracle):
    f, atlas = frames.quick_frame(frames.config(0), width=160, height=90)
    img = channels(oracle.rasterize(f), 160, 90)
    assert f.mesh.opaque.size and f.mesh.see_through.size
    assert len(np.unique(img.reshape(-1, 4), axis=0)) > 500
    assert np.all(img[..., 3] == 255)
