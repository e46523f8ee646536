"""The reference kernel tests' hand-built scenes, restated as numpy buffers.

crates/lumenwright-kernel/src/tests.rs builds these scenes in Rust (TestScene :37-126,
lamp_scene :310-395, the 16-brick wall :606-654, map_scene :725-741). They need no game
assets. Each builder returns a dict of buffers that ``view()`` turns into a ctypes
lw_scene_view for the oracle or the HIP library.
"""

from __future__ import annotations

import numpy as np

from lumenwright_amd.abi import (ALPHA_CUTOUT, ALPHA_OPAQUE, BRICK_WORDS, EMITTER,
                                 EMITTER_GRID_HEADER, EMPTY, MODEL, OV_CELL_WORDS, QUAD,
                                 UNIFORM, BoundView, f32_bits, new_params)


def quad(p0, eu, ev, flags):
    """tests.rs:13-22: a quad sampling a 2x1 atlas (texel 0 white, texel 1 clear)."""
    q = np.zeros((), dtype=QUAD)
    q["p0"] = [p0[0], p0[1], p0[2], f32_bits(flags)]
    q["eu"] = [eu[0], eu[1], eu[2], f32_bits(0xFFFFFFFF)]
    q["ev"] = [ev[0], ev[1], ev[2], 0.0]
    q["uv"] = [0.0, 0.0, 2.0, 0.0]
    q["uvt"] = [0.0, 1.0, 0.0, 0.0]
    return q


def cube(flags):
    """tests.rs:25-35: unit cube faces, outward winding (normal = ev x eu)."""
    x, y, z = np.array([1.0, 0, 0]), np.array([0, 1.0, 0]), np.array([0, 0, 1.0])
    return [
        quad((1, 1, 0), -x, -y, flags),  # north
        quad((0, 1, 1), x, -y, flags),   # south
        quad((1, 1, 1), -z, -y, flags),  # east
        quad((0, 1, 0), z, -y, flags),   # west
        quad((0, 1, 0), x, z, flags),    # up
        quad((0, 0, 1), x, -z, flags),   # down
    ]


def model(off, count):
    m = np.zeros((), dtype=MODEL)
    m["quad_offset"], m["quad_count"] = off, count
    return m


class Scene:
    def __init__(self, params, grid, models, quads, atlas, emitters, emitter_grid, sky):
        self.params = params
        self.grid = np.asarray(grid, dtype="<u4")
        self.models = np.asarray(models, dtype=MODEL)
        self.quads = np.asarray(quads, dtype=QUAD)
        self.atlas = np.asarray(atlas, dtype="<u4")
        self.emitters = np.asarray(emitters, dtype=EMITTER)
        self.emitter_grid = np.asarray(emitter_grid, dtype="<u4")
        self.sky = np.asarray(sky, dtype="<f4")

    def view(self):
        return BoundView(self.params, self.grid, self.models, self.quads, self.atlas,
                         self.emitters, self.emitter_grid, self.sky)


def test_scene(entity: bool) -> Scene:
    """tests.rs:48-112: 3x1x1 bricks -- brick 0 empty, brick 1 one cube at local (4,5,5),
    brick 2 uniform cubes; with `entity`, a cutout quad at x = 8.5 in brick 0's overlay."""
    quads = cube(ALPHA_OPAQUE)
    models = [model(0, 0), model(0, 6)]
    nb = 3
    top = [EMPTY, 0, UNIFORM | 1]
    brick = np.zeros(BRICK_WORDS, dtype="<u4")
    local = (5 * 16 + 5) * 16 + 4
    brick[local >> 1] |= 1 << ((local & 1) * 16)
    ov_top = [EMPTY] * nb
    ov_cells, ov_quads = [], []
    if entity:
        quads.append(quad((8.5, 6.0, 5.0), (0, 0, 1.0), (0, -1.0, 0), ALPHA_CUTOUT))
        ov_top[0] = 0
        ov_cells = np.zeros(OV_CELL_WORDS, dtype="<u4")
        cell = (5 * 16 + 5) * 16 + 8
        ov_cells[cell * 2] = 0
        ov_cells[cell * 2 + 1] = 1
        ov_quads = [6]
    grid = []
    off_top = len(grid); grid += top
    off_bricks = len(grid); grid += list(brick)
    off_ov_top = len(grid); grid += ov_top
    off_ov_cells = len(grid); grid += list(ov_cells)
    off_ov_quads = len(grid); grid += ov_quads
    grid.append(0)
    p = new_params()
    p["grid_dims"] = [3, 1, 1, nb]
    p["off"] = [off_top, off_bricks, off_ov_top, off_ov_cells]
    p["off2"] = [off_ov_quads, 2, 1, len(quads)]
    p["misc_f"] = [1.0, 1e-3, 0.0, 0.0]
    return Scene(p, grid, models, quads, [0xFFFFFFFF, 0x00000000], np.zeros(1, EMITTER),
                 np.zeros(EMITTER_GRID_HEADER, "<u4"), [])


def tower_scene(filled: bool = True) -> Scene:
    """3x6x3 bricks with three non-empty bricks in different rows and corners, for the walk
    that stops once every brick left toward the ray's octant is empty: a uniform brick of
    cubes at brick (1, 0, 1), an overlay-only brick at (0, 3, 2) holding a cutout quad in cell
    (8, 53, 37), and a stored brick at (2, 4, 0) with one cube in cell (36, 69, 5). Without
    `filled` every brick is EMPTY."""
    quads = cube(ALPHA_OPAQUE)
    models = [model(0, 0), model(0, 6)]
    dx, dy, dz = 3, 6, 3
    nb = dx * dy * dz
    brick_of = lambda bx, by, bz: (by * dz + bz) * dx + bx
    top = [EMPTY] * nb
    ov_top = [EMPTY] * nb
    brick = np.zeros(BRICK_WORDS, dtype="<u4")
    ov_cells, ov_quads = [], []
    if filled:
        top[brick_of(1, 0, 1)] = UNIFORM | 1
        top[brick_of(2, 4, 0)] = 0
        local = (5 * 16 + 5) * 16 + 4
        brick[local >> 1] |= 1 << ((local & 1) * 16)
        quads.append(quad((8.5, 54.0, 37.0), (0, 0, 1.0), (0, -1.0, 0), ALPHA_CUTOUT))
        ov_top[brick_of(0, 3, 2)] = 0
        ov_cells = np.zeros(OV_CELL_WORDS, dtype="<u4")
        cell = (5 * 16 + 5) * 16 + 8
        ov_cells[cell * 2 + 1] = 1
        ov_quads = [6]
    grid = []
    off_top = len(grid); grid += top
    off_bricks = len(grid); grid += list(brick)
    off_ov_top = len(grid); grid += ov_top
    off_ov_cells = len(grid); grid += list(ov_cells)
    off_ov_quads = len(grid); grid += ov_quads
    grid.append(0)
    p = new_params()
    p["grid_dims"] = [dx, dy, dz, nb]
    p["off"] = [off_top, off_bricks, off_ov_top, off_ov_cells]
    p["off2"] = [off_ov_quads, 2, 1, len(quads)]
    p["misc_f"] = [1.0, 1e-3, 0.0, 0.0]
    return Scene(p, grid, models, quads, [0xFFFFFFFF, 0x00000000], np.zeros(1, EMITTER),
                 np.zeros(EMITTER_GRID_HEADER, "<u4"), [])


LAMP_EMITTANCE = 0.5   # tests.rs:301
LAMP_INTENSITY = 13.0  # tests.rs:303


def lamp_scene(strategy: int, decoy: bool) -> Scene:
    """tests.rs:310-395: a 16-wide white floor at y = 0 and a lamp quad facing down at
    the bottom of block (7, 4, 7); with `decoy`, an up-facing lamp replaces the floor cube
    at (2, 0, 2), below the floor's top."""
    quads = cube(ALPHA_OPAQUE)
    lamp = quad((0, 0, 0), (0, 0, 1.0), (1.0, 0, 0), ALPHA_OPAQUE)
    lamp["ev"][3] = LAMP_EMITTANCE
    quads.append(lamp)
    up = quad((0, 0, 0), (1.0, 0, 0), (0, 0, 1.0), ALPHA_OPAQUE)
    up["ev"][3] = LAMP_EMITTANCE
    quads.append(up)
    for q in quads:
        q["uv"][2] = 1.0
    models = [model(0, 0), model(0, 6), model(6, 1), model(7, 1)]
    brick = np.zeros(BRICK_WORDS, dtype="<u4")

    def put(x, y, z, m):
        local = (y * 16 + z) * 16 + x
        brick[local >> 1] |= m << ((local & 1) * 16)

    for z in range(16):
        for x in range(16):
            if not (decoy and x == 2 and z == 2):
                put(x, 0, z, 1)
    put(7, 4, 7, 2)
    if decoy:
        put(2, 0, 2, 3)
    grid = [0]
    off_bricks = len(grid); grid += list(brick)
    off_ov_top = len(grid); grid.append(EMPTY)
    off_ov_cells = len(grid)
    off_ov_quads = len(grid)
    grid.append(0)
    em = [np.zeros((), EMITTER)]
    em[0]["pos"] = [7.0, 4.0, 7.0, f32_bits(6)]
    em[0]["faces"] = [1, 0, 0, 0]
    if decoy:
        e = np.zeros((), EMITTER)
        e["pos"] = [2.0, 0.0, 2.0, f32_bits(7)]
        e["faces"] = [1, 0, 0, 0]
        em.append(e)
    eg = [16, 1, 1, 1, 0, 0, 0, 0, 0, len(em)] + list(range(len(em)))
    p = new_params()
    p["grid_dims"] = [1, 1, 1, 1]
    p["off"] = [0, off_bricks, off_ov_top, off_ov_cells]
    p["off2"] = [off_ov_quads, 2, 1, len(quads)]
    p["misc_f"] = [LAMP_INTENSITY, 1e-3, 0.0, 0.0]
    p["misc_u"] = [7, 1, 4, len(em)]
    p["emitter"] = [strategy, 0, 0, 0]
    p["frame"] = [8, 8, 0, 3]
    p["sun_dir"] = [0.0, -1.0, 0.0, 1.0]
    return Scene(p, grid, models, quads, [0xFFFFFFFF, 0x00000000], em, eg, [])


def wall_scene() -> Scene:
    """tests.rs:606-654: 16x1x1 bricks; bricks 0-14 hold one corner cube, brick 15 a wall
    of front-only cubes at x = 250."""
    from lumenwright_amd.abi import FLAG_FRONT_ONLY
    quads = cube(ALPHA_OPAQUE | FLAG_FRONT_ONLY)
    models = [model(0, 0), model(0, 6)]
    nb = 16
    top = [0] * nb
    top[15] = 1
    stray = np.zeros(BRICK_WORDS, dtype="<u4")
    corner = (15 * 16 + 15) * 16
    stray[corner >> 1] |= 1 << ((corner & 1) * 16)
    wall = np.zeros(BRICK_WORDS, dtype="<u4")
    for y in range(16):
        for z in range(16):
            local = (y * 16 + z) * 16 + (250 - 240)
            wall[local >> 1] |= 1 << ((local & 1) * 16)
    grid = list(top)
    off_bricks = len(grid)
    grid += list(stray) + list(wall)
    off_ov_top = len(grid)
    grid += [EMPTY] * nb
    off_ov_cells = len(grid)
    grid.append(0)
    p = new_params()
    p["grid_dims"] = [16, 1, 1, nb]
    p["off"] = [0, off_bricks, off_ov_top, off_ov_cells]
    p["off2"] = [off_ov_cells, 2, 1, len(quads)]
    return Scene(p, grid, models, quads, [0xFFFFFFFF, 0x00000000], np.zeros(1, EMITTER),
                 np.zeros(EMITTER_GRID_HEADER, "<u4"), [])


def wall_rays():
    """tests.rs:658-670: 4096 rays crossing ~250 cells towards the wall."""
    o, d = [], []
    for i in range(64):
        for j in range(64):
            o.append([np.float32(0.137) + np.float32(i) * np.float32(0.0131),
                      np.float32(1.1) + np.float32(i) * np.float32(0.03),
                      np.float32(1.3) + np.float32(j) * np.float32(0.05)])
            d.append([np.float32(1.0),
                      np.float32(0.04) + np.float32(0.002) * (np.float32(j) - np.float32(32.0))
                      / np.float32(32.0),
                      np.float32(0.03) + np.float32(0.002) * (np.float32(i) - np.float32(32.0))
                      / np.float32(32.0)])
    return np.array(o, dtype="<f4"), np.array(d, dtype="<f4")


def map_scene() -> Scene:
    """tests.rs:725-741: two sky maps -- the lens's 4x2 (texel (i, j) = (i, j, 1)) and the
    light's 2x2 ((10 + i, 10 + j, 2))."""
    s = test_scene(False)
    s.params["sky_map"] = [4.0, 2.0, 2.0, 2.0]
    sky = []
    for j in range(2):
        for i in range(4):
            sky += [i, j, 1.0]
    for j in range(2):
        for i in range(2):
            sky += [10.0 + i, 10.0 + j, 2.0]
    s.sky = np.array(sky, dtype="<f4")
    return s


# tests.rs:590-600: Preetham::new for a sun at altitude 45, azimuth 45 degrees.
PREETHAM_45_ZENITH = [0.29624206, 0.30751193, 8.306274, 0.0]
PREETHAM_PEREZ = [
    [-0.30745, -0.30255, -1.01625, 0.0],
    [-0.16545, -0.2283, -0.461, 0.0],
    [0.2115, 0.19045, 5.26835, 0.0],
    [-1.05915, -1.76395, -2.2756, 0.0],
    [0.03695, 0.02565, 0.2028, 0.0],
]


def normalize32(v):
    """math.rs:275-278 in f32: v * (1 / |v|)."""
    v = np.asarray(v, dtype=np.float32)
    d = np.float32(v[0] * v[0] + v[1] * v[1]) + np.float32(v[2] * v[2])
    l = np.sqrt(np.float32(d), dtype=np.float32)
    return (v * (np.float32(1.0) / l)).astype(np.float32)


def floor_scene(sun_altitude: float, sun_azimuth: float, sun_radius: float, width: int = 32,
                height: int = 32):
    """A white 64 x 64 floor (the top of a layer of cubes at y = 0) under the sun alone: no
    sky light (a sky cache of zeros), one bounce, a camera 11 blocks above looking down.
    The image is sun_color * E[max(n.l, 0)] over the sun's cone, a closed form:
    (1 + cos a) / 2 * sin(altitude) for a cone of half-angle a wholly above the floor's
    horizon. Returns the Scene and d sun_dir / d(altitude, azimuth)."""
    from lumenwright_amd.abi import FLAG_FRONT_ONLY
    quads = cube(ALPHA_OPAQUE | FLAG_FRONT_ONLY)
    for q in quads:
        q["uv"][2] = 1.0  # texel 0 only: white
    models = [model(0, 0), model(0, 6)]
    brick = np.zeros(BRICK_WORDS, dtype="<u4")
    for z in range(16):
        for x in range(16):
            local = z * 16 + x
            brick[local >> 1] |= 1 << ((local & 1) * 16)
    nb = 16
    grid = [0] * nb
    off_bricks = len(grid); grid += list(brick)
    off_ov_top = len(grid); grid += [EMPTY] * nb
    off_ov_cells = len(grid)
    grid.append(0)
    alt, az = np.float64(sun_altitude), np.float64(sun_azimuth)
    s = np.array([np.sin(az) * np.cos(alt), np.sin(alt), -np.cos(az) * np.cos(alt)])
    d_alt = np.array([-np.sin(az) * np.sin(alt), np.cos(alt), np.cos(az) * np.sin(alt)])
    d_az = np.array([np.cos(az) * np.cos(alt), 0.0, np.sin(az) * np.cos(alt)])
    n_cache = 4
    p = new_params()
    p["grid_dims"] = [4, 1, 4, nb]
    p["off"] = [0, off_bricks, off_ov_top, off_ov_cells]
    p["off2"] = [off_ov_cells, 2, 1, len(quads)]
    p["misc_f"] = [0.0, 1e-3, float(n_cache), 0.0]
    p["misc_u"] = [3, 1, 2, 0]
    p["frame"] = [width, height, 0, 1]
    p["cam_pos"] = [32.0, 12.0, 32.0, 0.4]
    p["cam_fwd"] = [0.0, -1.0, 0.0, width / height]
    p["cam_right"] = [1.0, 0.0, 0.0, 0.0]
    p["cam_up"] = [0.0, 0.0, -1.0, 0.0]
    p["sun_dir"] = [s[0], s[1], s[2], np.cos(np.float64(sun_radius))]
    p["sun_color"] = [1.0, 1.0, 1.0, 0.0]
    p["sky_sun"] = [s[0], s[1], s[2], 1.0]
    sky = np.zeros((n_cache + 1) ** 2 * 3, dtype="<f4")
    scene = Scene(p, grid, models, quads, [0xFFFFFFFF, 0x00000000], np.zeros(1, EMITTER),
                  np.zeros(EMITTER_GRID_HEADER, "<u4"), sky)
    return scene, np.stack([d_alt, d_az])
