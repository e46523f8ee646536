"""The walk's early exit, on the CPU: a ray stops once every brick left toward its octant is
EMPTY with no overlay. The oracle's counted build stops its walk there as the device does.

* The oracle's per-brick octant bits equal the box they stand for, taken by numpy slicing of
  the grid's occupied bricks.
* The stopped walk renders the same image, bit for bit, as the whole walk of the reference
  (trace.rs:137-290): the exit changes where a walk that hits nothing ends, never whether
  it hits.
"""

import numpy as np
import pytest

import scenes
from lumenwright_amd import frames
from lumenwright_amd.abi import EMPTY


def occupied(view):
    p = view.params
    dx, dy, dz = (int(x) for x in p["grid_dims"][:3])
    nb = dx * dy * dz
    g = view.grid
    top = g[int(p["off"][0]):int(p["off"][0]) + nb]
    ov = g[int(p["off"][2]):int(p["off"][2]) + nb]
    return ((top != EMPTY) | (ov != EMPTY)).reshape(dy, dz, dx)


def boxes_by_slicing(occ):
    dy, dz, dx = occ.shape
    out = np.zeros(occ.shape, dtype="u1")
    for y in range(dy):
        for z in range(dz):
            for x in range(dx):
                for k in range(8):
                    sx = slice(x, None) if k & 1 else slice(0, x + 1)
                    sy = slice(y, None) if k & 2 else slice(0, y + 1)
                    sz = slice(z, None) if k & 4 else slice(0, z + 1)
                    if occ[sy, sz, sx].any():
                        out[y, z, x] |= 1 << k
    return out.reshape(-1)


@pytest.mark.parametrize("which", ["tower", "tower_empty", "test_scene", "cfg0"])
def test_the_octant_bits_are_the_boxes(oracle, which):
    if which == "cfg0":
        v = frames.make_frame(frames.config(0, width=32, height=18)).view()
    elif which == "test_scene":
        v = scenes.test_scene(False).view()  # 3 x 1 x 1 bricks, the first one EMPTY
    else:
        v = scenes.tower_scene(which == "tower").view()
    occ = occupied(v)
    got = oracle.octant_boxes(v)
    want = boxes_by_slicing(occ)
    assert np.array_equal(got, want)
    if which == "tower_empty":
        assert not got.any()
    else:
        assert got.any() and (got != 0xFF).any()  # some octants open, some closed


@pytest.mark.parametrize("which,size,spp", [(0, (96, 54), 4), (1, (64, 36), 2)])
def test_the_stopped_walk_renders_the_same_image(oracle, which, size, spp):
    cfg = frames.config(which, width=size[0], height=size[1])
    v = frames.make_frame(cfg).view()
    whole = oracle.render_hq(v, spp)
    stopped, c = oracle.render_hq_counted(v, spp)
    assert c.samples == size[0] * size[1] * spp
    assert np.array_equal(whole.view("<u4"), stopped.view("<u4"))
