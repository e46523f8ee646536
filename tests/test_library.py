"""The C-ABI library loads, exports every function include/lw/*.h declares, and reports a
missing device as an error instead of failing (render/cuda.rs:29-41); no compute calls."""

import ctypes
import re
from pathlib import Path

import numpy as np

import lumenwright_amd as lw
from lumenwright_amd import _native

ROOT = Path(__file__).resolve().parent.parent


def declared_functions():
    names = []
    for h in sorted((ROOT / "include" / "lw").glob("*.h")):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        names += re.findall(r"\b(lw_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def test_every_declared_symbol_is_exported():
    L = _native.lib()
    names = declared_functions()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # and the ctypes signature table covers each of them
    assert set(names) <= set(_native.SIGNATURES), set(names) - set(_native.SIGNATURES)


def test_abi_struct_sizes_match_the_reference():
    # trace.rs:327-333; lw_hit and lw_timing are this library's own.
    assert lw.PARAMS.itemsize == 432 and lw.QUAD.itemsize == 80
    assert lw.MODEL.itemsize == 16 and lw.EMITTER.itemsize == 32 and lw.HIT.itemsize == 40
    from lumenwright_amd.abi import TimingC
    assert ctypes.sizeof(TimingC) == 40


def test_no_device_is_an_error_not_a_crash(monkeypatch):
    import os
    # On a machine with a GPU this test only checks the error path of a bad index.
    try:
        desc = lw.probe()
    except lw.BackendError as e:
        assert e.code == -2 and "no HIP device" in str(e)
        try:
            lw.HipBackend.open(0)
            raise AssertionError("open succeeded without a device")
        except lw.BackendError as e2:
            assert e2.code == -2
        return
    assert desc
    try:
        lw.HipBackend.open(1 << 20)
        raise AssertionError("open of a device that does not exist succeeded")
    except lw.BackendError as e:
        assert e.code == -1


def test_oracle_cpu_backend_is_thread_count_invariant(oracle):
    # library.rs:400-478: bit-identical HDR radiance on 1, 3 and 8 threads.
    from lumenwright_amd import frames
    fr = frames.make_frame(frames.config(1, width=64, height=36, spp=4, world_chunks_x=4,
                                         world_chunks_z=4))
    v = fr.view()
    imgs = [oracle.render_hq(v, 4, threads=t) for t in (1, 3, 8)]
    assert all(np.array_equal(imgs[0].view("<u4"), im.view("<u4")) for im in imgs[1:])
    assert np.isfinite(imgs[0]).all() and imgs[0].mean() > 0


def test_per_view_arrays_are_checked_before_the_call():
    # lw_hip_adjoint_batch and lw_hip_fit_new copy w*h*3 floats from each view's pointer, so
    # the binding refuses an array of another size before it reaches the library.
    import numpy as np
    import pytest
    from lumenwright_amd.abi import new_params
    from lumenwright_amd.backend import HipBackend
    p = new_params()
    p["frame"][:2] = [4, 3]
    views = [(p, 1), (p, 2)]
    ok = HipBackend._per_view_images(views, [np.zeros(36), np.zeros((12, 3))], "dL/dI")
    assert [a.dtype.str for a in ok] == ["<f4", "<f4"] and all(a.size == 36 for a in ok)
    with pytest.raises(ValueError, match="view 1"):
        HipBackend._per_view_images(views, [np.zeros(36), np.zeros(35)], "dL/dI")
    with pytest.raises(ValueError, match="2 views"):
        HipBackend._per_view_images(views, [np.zeros(36)], "target")
