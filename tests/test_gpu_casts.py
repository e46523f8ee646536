"""The device's f32 -> integer conversions (lw_hip_casts, the helper every kernel uses for
Rust's `as i32` / `as u32`, e.g. trace.rs:161-163, scene.rs:291, integrator.rs:351) against
Rust's defined semantics, restated here with numpy: truncation toward zero, saturation at the
target type's range, NaN to 0.

Bar: bit-exact on special values (signed zeros, infinities, NaNs with any payload and sign,
subnormals, the neighbours of every saturation bound) and on 2^20 random bit patterns."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rust_as_i32(x):
    with np.errstate(invalid="ignore"):
        x = x.astype("<f8")
    out = np.zeros(x.shape, dtype="<i8")
    ok = ~np.isnan(x)
    out[ok] = np.trunc(np.clip(x[ok], -2147483648.0, 2147483647.0)).astype("<i8")
    return out.astype("<i4")


def rust_as_u32(x):
    with np.errstate(invalid="ignore"):
        x = x.astype("<f8")
    out = np.zeros(x.shape, dtype="<i8")
    ok = ~np.isnan(x)
    out[ok] = np.trunc(np.clip(x[ok], 0.0, 4294967295.0)).astype("<i8")
    return out.astype("<u4")


def specials():
    f = np.float32
    vals = [0.0, -0.0, np.inf, -np.inf, 0.5, -0.5, 0.9999999, -0.9999999, 1.0, -1.0, 1.5, -1.5,
            2.5, -2.5, 8388607.5, 16777217.0, np.finfo(f).max, -np.finfo(f).max,
            np.finfo(f).tiny, -np.finfo(f).tiny]
    for b in (2.0 ** 31, -2.0 ** 31, 2.0 ** 32, 2.0 ** 24):
        c = f(b)
        vals += [c, np.nextafter(c, f(np.inf)), np.nextafter(c, f(-np.inf))]
    x = np.array(vals, dtype="<f4")
    bits = np.array([0x7FC00000, 0xFFC00000, 0x7F800001, 0xFF800001, 0x7FFFFFFF, 0xFFFFFFFF,
                     0x00000001, 0x80000001, 0x007FFFFF, 0x807FFFFF], dtype="<u4").view("<f4")
    return np.concatenate([x, bits])


def test_casts_follow_rust_semantics(hip_backend):
    rng = np.random.default_rng(5)
    x = np.concatenate([specials(), rng.integers(0, 2 ** 32, 1 << 20, dtype=np.uint64)
                        .astype("<u4").view("<f4"),
                        rng.uniform(-5e9, 5e9, 1 << 16).astype("<f4")])
    gi, gu = hip_backend.casts(x)
    wi, wu = rust_as_i32(x), rust_as_u32(x)
    bad_i = np.flatnonzero(gi != wi)
    bad_u = np.flatnonzero(gu != wu)
    print(f"{x.size} values; as i32 differs at {bad_i.size}, as u32 at {bad_u.size}")
    assert bad_i.size == 0, [(x[k], gi[k], wi[k]) for k in bad_i[:8]]
    assert bad_u.size == 0, [(x[k], gu[k], wu[k]) for k in bad_u[:8]]
