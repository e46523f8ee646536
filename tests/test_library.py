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


def _functions(text):
    """(name, params, body, is_kernel, head) of each function definition in HIP/C++ source
    text (comments blanked), found by its parameter list and matched braces; head is the text
    before the name back to the previous statement or block."""
    text = re.sub(r"//[^\n]*", "", text)
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    out = []
    for m in re.finditer(r"\b([A-Za-z_]\w*)\s*\(", text):
        name = m.group(1)
        if name in ("if", "for", "while", "switch", "return", "sizeof", "static_assert"):
            continue
        depth, i = 1, m.end()
        while i < len(text) and depth:
            depth += {"(": 1, ")": -1}.get(text[i], 0)
            i += 1
        params = text[m.end():i - 1]
        j = i
        while j < len(text) and text[j] in " \t\n":
            j += 1
        while text.startswith("const", j) or text.startswith("noexcept", j):
            j = text.index(" ", j) if " " in text[j:j + 9] else j + 5
            while j < len(text) and text[j] in " \t\n":
                j += 1
        if j >= len(text) or text[j] != "{":
            continue
        depth, k = 1, j + 1
        while k < len(text) and depth:
            depth += {"{": 1, "}": -1}.get(text[k], 0)
            k += 1
        head = text[max(0, m.start() - 300):m.start()].split(";")[-1].split("}")[-1]
        out.append((name, params, text[j + 1:k - 1], "__global__" in head, head))
    return out


def test_every_kernel_that_reads_its_arguments_again_takes_params_and_scene_first():
    # The walk and the event loop read the kernel's (p, s) again from its kernel argument
    # segment (kernel_args in lw_device.hpp), at offset 0 and right after lw_params. Every
    # kernel that can reach that read must take (const lw_params p, const DevScene s) first.
    csrc = ROOT / "lumenwright_amd" / "csrc"
    funcs = []
    for f in sorted(csrc.glob("*.hpp")) + sorted(csrc.glob("*.hip")):
        funcs += _functions(f.read_text())
    calls = {}
    for name, _, body, _, _ in funcs:
        calls.setdefault(name, set()).update(re.findall(r"\b([A-Za-z_]\w*)\s*[(<]", body))
    reach = {"kernel_args"}
    while True:
        more = {n for n, c in calls.items() if n not in reach and c & reach}
        if not more:
            break
        reach |= more
    assert {"walk_step", "walk_begin", "trace", "sample_event", "render_sample"} <= reach
    kernels = [(n, p, b) for n, p, b, g, _ in funcs if g]
    assert len(kernels) >= 30
    bad = []
    checked = 0
    for name, params, body in kernels:
        if set(re.findall(r"\b([A-Za-z_]\w*)\s*[(<]", body)) & reach:
            checked += 1
            first = [x.strip() for x in params.split(",")[:2]]
            if first != ["const lw_params p", "const DevScene s"]:
                bad.append((name, first))
    assert checked >= 12 and not bad, bad
    # A called (not inlined) function has no kernel argument segment of its own to read, so
    # every device function on the way to kernel_args is force-inlined into its kernel.
    not_inlined = sorted({n for n, _, _, g, head in funcs
                          if n in reach and not g and "__device__" in head and "__forceinline__" not in head})
    assert not not_inlined, not_inlined
