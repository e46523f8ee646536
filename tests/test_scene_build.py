"""build_hq's layout (crates/lumenwright/src/scene/voxel.rs:305-491, emitter_list
voxel.rs:542-607) restated by the oracle, pinned against the reference's hand-built kernel
test scene (crates/lumenwright-kernel/src/tests.rs:48-112) and its tests, the emitter-grid
test of voxel.rs:616-671, decoding every voxel back out of the grid, and brute-force overlay
lists; then the host world builder's layout against the oracle's."""

import dataclasses

import numpy as np
import pytest

import scenes
from lumenwright_amd import frames
from lumenwright_amd.abi import (ALPHA_CUTOUT, ALPHA_OPAQUE, EMITTER_GRID_HEADER, EMPTY,
                                 MODEL, MODEL_EMITS, MODEL_OCCLUDES, OV_CELL_WORDS, QUAD,
                                 UNIFORM, BoundView, VoxelVolume, f32_bits, new_params)


def kernel_test_volume(entity: bool) -> VoxelVolume:
    """tests.rs:48-112 as an interned volume: brick 0 empty, brick 1 one cube at local
    (4, 5, 5), brick 2 all cubes; with `entity`, the cutout quad at x = 8.5."""
    vox = np.zeros(3 * 4096, dtype="<u2")
    vox[4096 + (5 * 16 + 5) * 16 + 4] = 1
    vox[2 * 4096:] = 1
    ent = [scenes.quad((8.5, 6.0, 5.0), (0, 0, 1.0), (0, -1.0, 0), ALPHA_CUTOUT)] if entity else []
    return VoxelVolume(dims=(3, 1, 1), voxels=vox,
                       models=np.array([scenes.model(0, 0), scenes.model(0, 6)], dtype=MODEL),
                       model_flags=np.array([0, MODEL_OCCLUDES], np.uint8),
                       quads=np.array(scenes.cube(ALPHA_OPAQUE), dtype=QUAD),
                       atlas=np.array([0xFFFFFFFF, 0], "<u4"), atlas_size=(2, 1),
                       entity_quads=np.array(ent, dtype=QUAD))


def view_of(vol: VoxelVolume, built, sky=None, params=None) -> BoundView:
    """The oracle-built scene as a view, its Params carrying the built layout."""
    grid, em, eg, info = built
    nb = int(np.prod(vol.dims))
    p = new_params() if params is None else params.copy()
    quads = np.concatenate([vol.quads, vol.entity_quads]) if len(vol.entity_quads) else vol.quads
    p["grid_dims"] = [*vol.dims, nb]
    p["off"] = info["offsets"][:4]
    p["off2"] = [info["offsets"][4], vol.atlas_size[0], vol.atlas_size[1], len(quads)]
    p["see_through"] = [*info["see_through"], 0, 0]
    if params is None:
        p["misc_f"] = [1.0, 1e-3, 0.0, 0.0]
    p["misc_u"][2] = len(vol.models)
    p["misc_u"][3] = info["emitter_count"]
    return BoundView(p, grid, vol.models, quads, vol.atlas, em, eg,
                     np.zeros(0, "<f4") if sky is None else sky)


def test_the_layout_of_the_kernel_test_scene_is_the_reference_layout(oracle):
    # tests.rs:48-112 lays the scene out by hand: the built grid is that buffer word for word.
    vol = kernel_test_volume(False)
    grid, em, eg, info = oracle.build_layout(vol)
    hand = scenes.test_scene(False)
    assert np.array_equal(grid, hand.grid)
    assert list(info["offsets"][:4]) == list(hand.params["off"])
    assert info["offsets"][4] == hand.params["off2"][0]
    assert info["uniform_bricks"] == 1 and info["bricks"] == 1 and info["buried_bricks"] == 0
    # no emitters: one unused entry and the bare header (voxel.rs:543-547)
    assert info["emitter_count"] == 0 and len(em) == 1 and not em.view("u1").any()
    assert list(eg) == [10, 0, 0, 0, 0, 0, 0, 0]


def one(o, d, view, oracle, tmin=0.0, tmax=np.inf):
    h = oracle.trace(view, np.array([o], "<f4"), np.array([d], "<f4"), tmin, tmax)[0]
    return h if h["hit"] else None


def test_the_kernel_tests_pass_on_the_built_scene(oracle):
    # tests.rs:128-222 on the built scene (with the entity overlay build_hq makes)
    vol = kernel_test_volume(True)
    v = view_of(vol, oracle.build_layout(vol))
    h = one([0.5, 5.5, 5.5], [1, 0, 0], v, oracle)
    assert abs(h["t"] - 19.5) < 1e-3 and h["quad"] == 3 and list(h["cell"]) == [20, 5, 5]
    assert abs(one([-10.0, 5.5, 5.5], [1, 0, 0], v, oracle)["t"] - 30.0) < 1e-3
    assert abs(one([0.5, 9.5, 5.5], [1, 0, 0], v, oracle)["t"] - 31.5) < 1e-3
    assert one([0.5, 5.5, 5.5], [0, 1, 0], v, oracle) is None
    h = one([0.5, 5.5, 5.25], [1, 0, 0], v, oracle)
    assert abs(h["t"] - 8.0) < 1e-3 and h["quad"] == 6
    h = one([0.5, 5.5, 5.75], [1, 0, 0], v, oracle)
    assert abs(h["t"] - 19.5) < 1e-3


def emitter_volume(sites, dims):
    vox = np.zeros(int(np.prod(dims)) * 4096, dtype="<u2")
    dx, dy, dz = dims
    for x, y, z in sites:
        b = ((y >> 4) * dz + (z >> 4)) * dx + (x >> 4)
        vox[b * 4096 + ((y & 15) * 16 + (z & 15)) * 16 + (x & 15)] = 1
    models = np.zeros(2, dtype=MODEL)
    models[1]["quad_offset"], models[1]["quad_count"] = 3, 2
    quads = np.zeros(5, dtype=QUAD)
    return VoxelVolume(dims=dims, voxels=vox, models=models,
                       model_flags=np.array([0, MODEL_OCCLUDES | MODEL_EMITS], np.uint8),
                       quads=quads, atlas=np.array([0xFFFFFFFF], "<u4"), atlas_size=(1, 1))


def test_the_emitter_grid_lists_every_emitter_within_a_cell(oracle):
    # voxel.rs:616-671: from any point, every emitter less than a cell away along each axis is
    # a candidate and none is two cells or more away, read back through the kernel's lookup.
    sites = [(25, 3, 0), (4, 40, 17), (26, 3, 1)]
    vol = emitter_volume(sites, (2, 3, 2))
    built = oracle.build_layout(vol)
    grid, em, eg, info = built
    assert info["emitter_count"] == 3 and len(em) == 3
    # the list follows brick then voxel order: (25,3,0) and (26,3,1) share brick 1
    pos = [tuple(int(c) for c in e["pos"][:3]) for e in em]
    assert pos == [(25, 3, 0), (26, 3, 1), (4, 40, 17)]
    assert list(em[2]["pos"].view("<u4"))[3] == 3 and em[2]["faces"][0] == 2
    v = view_of(vol, built)
    cells = int(eg[1]) * int(eg[2]) * int(eg[3])
    lst = eg[EMITTER_GRID_HEADER + 2 * cells:]
    p = np.stack(np.meshgrid(np.arange(-5, 50), np.arange(-5, 60), np.arange(-5, 40),
                             indexing="ij"), -1).reshape(-1, 3)
    checked = 0
    for x, y, z in p:
        start, count = oracle.emitter_cell(v, [x + 0.5, y + 0.5, z + 0.5])
        found = set(lst[start:start + count].tolist())
        for i, c in enumerate(pos):
            d = np.abs(np.array([x, y, z]) - c)
            if np.all(d < 10):
                assert i in found, (i, x, y, z)
            if np.any(d >= 20):
                assert i not in found, (i, x, y, z)
            checked += 1
    assert checked == 55 * 65 * 45 * 3


def decode(grid, offsets, dims, x, y, z):
    """The model id the kernel reads at voxel (x, y, z) (scene.rs:1-9 layout)."""
    dx, _, dz = dims
    e = int(grid[offsets[0] + ((y >> 4) * dz + (z >> 4)) * dx + (x >> 4)])
    if e == EMPTY:
        return 0
    if e & UNIFORM:
        return e & ~UNIFORM
    local = ((y & 15) * 16 + (z & 15)) * 16 + (x & 15)
    w = int(grid[offsets[1] + e * 2048 + (local >> 1)])
    return (w >> ((local & 1) * 16)) & 0xFFFF


DEEP = frames.WorldSpec(3, 3, 8, cube_types=1, water_level=62, emitter_fraction=0.02,
                        cutout_types=2, cutout_fraction=0.2, seed=5)


@pytest.mark.parametrize("given_occlusion", [True, False])
def test_the_layout_decodes_to_the_volume(oracle, given_occlusion):
    # Every voxel of a non-buried brick reads back its id; a buried brick reads its first set
    # voxel's id and all six neighbours occlude; EMPTY exactly where nothing is set; the
    # emitter list is every emitting voxel in brick then voxel order.
    scene = frames.build_world(DEEP, volume=True)
    vol = scene.volume if given_occlusion else dataclasses.replace(scene.volume, brick_occludes=None)
    grid, em, eg, info = oracle.build_layout(vol)
    dx, dy, dz = vol.dims
    nb = dx * dy * dz
    ids = vol.voxels.reshape(nb, 4096)
    occl = np.all(vol.model_flags[ids] & MODEL_OCCLUDES, axis=1)
    if given_occlusion:
        assert np.array_equal(occl, vol.brick_occludes.astype(bool))
    top = grid[:nb]
    assert info["buried_bricks"] > 0 and info["uniform_bricks"] > 0 and info["bricks"] > 0
    for b in range(nb):
        bx, by, bz = b % dx, b // (dx * dz), (b // dx) % dz
        if not ids[b].any():
            assert top[b] == EMPTY
            continue
        nbrs = [(bx + 1, by, bz), (bx - 1, by, bz), (bx, by + 1, bz), (bx, by - 1, bz),
                (bx, by, bz + 1), (bx, by, bz - 1)]
        inside = [0 <= a < dx and 0 <= c < dy and 0 <= e < dz for a, c, e in nbrs]
        buried = occl[b] and all(inside) and all(occl[(c * dz + e) * dx + a] for a, c, e in nbrs)
        if buried:
            assert top[b] == UNIFORM | int(ids[b][np.flatnonzero(ids[b])[0]])
        elif (ids[b] == ids[b][0]).all():
            assert top[b] == UNIFORM | int(ids[b][0])
        else:
            assert top[b] & UNIFORM == 0
            words = grid[info["offsets"][1] + int(top[b]) * 2048:][:2048]
            assert np.array_equal(words.view("<u2"), ids[b])
    assert info["bricks"] == int(np.sum((top & UNIFORM) == 0))  # EMPTY has the UNIFORM bit
    # a random sample of voxels through the kernel's addressing
    rng = np.random.default_rng(4)
    for _ in range(3000):
        x, y, z = (int(rng.integers(0, 16 * d)) for d in (dx, dy, dz))
        b = ((y >> 4) * dz + (z >> 4)) * dx + (x >> 4)
        if top[b] & UNIFORM and top[b] != EMPTY and (ids[b] != ids[b][0]).any():
            continue  # buried
        want = int(ids[b][((y & 15) * 16 + (z & 15)) * 16 + (x & 15)])
        assert decode(grid, info["offsets"], vol.dims, x, y, z) == want
    # emitters: brute force over the volume
    sites = []
    for b in range(nb):
        bx, by, bz = b % dx, b // (dx * dz), (b // dx) % dz
        for local in np.flatnonzero(vol.model_flags[ids[b]] & MODEL_EMITS):
            sites.append((bx * 16 + (local & 15), by * 16 + (local >> 8), bz * 16 + ((local >> 4) & 15),
                          int(ids[b][local])))
    assert info["emitter_count"] == len(sites) > 0
    assert [tuple(int(c) for c in e["pos"][:3]) for e in em] == [s[:3] for s in sites]
    first = vol.models["quad_offset"][[s[3] for s in sites]]
    assert np.array_equal(em["pos"][:, 3].view("<u4"), first)
    assert np.array_equal(em["faces"][:, 0], vol.models["quad_count"][[s[3] for s in sites]])


def random_entity_quads(rng, n, size):
    q = np.zeros(n, dtype=QUAD)
    for i in range(n):
        kind = i % 5
        p0 = rng.uniform(-4, np.array(size) + 4).astype("<f4")
        if kind == 0:    # a small quad inside one voxel or across a few
            eu, ev = rng.uniform(-1.5, 1.5, 3), rng.uniform(-1.5, 1.5, 3)
        elif kind == 1:  # a long thin one across bricks
            eu, ev = np.array([rng.uniform(10, 40), 0, 0]), np.array([0, rng.uniform(0.1, 2), 0])
        elif kind == 2:  # axis-aligned face on a voxel boundary
            p0 = np.floor(p0)
            eu, ev = np.array([0, 0, 1.0]), np.array([0, -1.0, 0])
        elif kind == 3:  # outside the grid
            p0 = np.array([-20.0, -20.0, -20.0]) - rng.uniform(0, 5, 3)
            eu, ev = np.array([1.0, 0, 0]), np.array([0, 1.0, 0])
        else:
            eu, ev = rng.uniform(-3, 3, 3), rng.uniform(-3, 3, 3)
        q[i]["p0"][:3], q[i]["eu"][:3], q[i]["ev"][:3] = p0, eu, ev
        q[i]["p0"][3] = f32_bits(ALPHA_CUTOUT)
    q[n - 1]["p0"][0] = np.nan  # NaN corners fold away (f32::min keeps the other operand)
    return q


def as_i32(f):
    """Rust `as i32` of floats: saturating, NaN -> 0."""
    f = np.asarray(f, dtype=np.float64)
    return np.where(np.isnan(f), 0, np.clip(np.nan_to_num(f), -2.0**31, 2.0**31 - 1)).astype(np.int64)


def overlay_brute_force(q, dims):
    """Per voxel cell (brick, local), the quads whose grown bounds touch it (voxel.rs:390-423)."""
    size = np.array(dims) * 16
    cells = {}
    for i in range(len(q)):
        p0, eu, ev = (q[i][k][:3].astype("<f4") for k in ("p0", "eu", "ev"))
        c = np.stack([p0, p0 + ev, (p0 + eu) + ev, p0 + eu]).astype("<f4")
        lo = np.fmin.reduce(np.vstack([np.full(3, 3.4028235e38, "<f4"), c]), axis=0) - np.float32(1e-3)
        hi = np.fmax.reduce(np.vstack([np.full(3, -3.4028235e38, "<f4"), c]), axis=0) + np.float32(1e-3)
        lo = np.maximum(as_i32(np.floor(lo.astype("<f4"))), 0)
        hi = np.minimum(as_i32(np.floor(hi.astype("<f4"))), size - 1)
        for y in range(lo[1], hi[1] + 1):
            for z in range(lo[2], hi[2] + 1):
                for x in range(lo[0], hi[0] + 1):
                    b = ((y >> 4) * dims[2] + (z >> 4)) * dims[0] + (x >> 4)
                    local = ((y & 15) * 16 + (z & 15)) * 16 + (x & 15)
                    cells.setdefault((b, local), []).append(i)
    return cells


def test_the_overlay_lists_every_quad_in_every_cell_it_touches(oracle):
    rng = np.random.default_rng(9)
    dims = (3, 2, 2)
    vol = dataclasses.replace(emitter_volume([(5, 5, 5), (40, 20, 30)], dims),
                              entity_quads=random_entity_quads(rng, 60, np.array(dims) * 16),
                              see_through_count=7)
    grid, em, eg, info = oracle.build_layout(vol)
    want = overlay_brute_force(vol.entity_quads, dims)
    base = len(vol.quads)
    off = info["offsets"]
    nb = int(np.prod(dims))
    ov_top = grid[off[2]:off[2] + nb]
    bricks = sorted({b for b, _ in want})
    assert [int(ov_top[b]) for b in bricks] == list(range(len(bricks)))  # ascending bricks
    assert all(ov_top[b] == EMPTY for b in range(nb) if b not in bricks)
    assert info["overlay_bricks"] == len(bricks)
    ov_quads = grid[off[4]:-1]
    seen = 0
    for b in bricks:
        cells = grid[off[3] + int(ov_top[b]) * OV_CELL_WORDS:][:OV_CELL_WORDS].reshape(4096, 2)
        for local in range(4096):
            start, count = int(cells[local, 0]), int(cells[local, 1])
            got = [int(v) - base for v in ov_quads[start:start + count]]
            assert got == want.get((b, local), []), (b, local)
            seen += count
    assert seen == len(ov_quads) == sum(len(v) for v in want.values())
    assert info["see_through"] == (base + 60 - 7, 7)
    assert grid[-1] == 0


@pytest.mark.parametrize("spec", [
    frames.WorldSpec(2, 2, 8, cube_types=4, water_level=62, emitter_fraction=0.02),
    DEEP,
    frames.WorldSpec(4, 4, 6, cube_types=8, emitter_blocks=12, seed=3),
], ids=["small", "deep", "emitter-blocks"])
def test_the_world_builders_layout_is_the_oracles(oracle, spec):
    scene = frames.build_world(spec, volume=True)
    grid, em, eg, info = oracle.build_layout(scene.volume)
    assert np.array_equal(grid, scene.grid)
    assert np.array_equal(em.view("u1"), scene.emitters.view("u1"))
    assert np.array_equal(eg, scene.emitter_grid)
    assert info["offsets"] == tuple(scene.offsets) and info["emitter_count"] == scene.emitter_count
    assert info["bricks"] == scene.stats["bricks"]
    assert info["uniform_bricks"] == scene.stats["uniform_bricks"]
    assert info["buried_bricks"] == scene.stats["buried_bricks"]
