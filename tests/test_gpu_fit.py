"""The batched inverse fit (config 4) on the device: its gradient against the CPU reference
(oracle/lw_oracle_diff.hpp) and against central differences, on a world of config 4's kind
(water, cutout plants, emitters sampled with ONE, the game look's dusk sky maps), and the fit
at config 4's full size.

Bars: lw_hip_adjoint and lw_hip_adjoint_batch equal the CPU reference's gradient within 1e-6
of its largest component (the same samples and estimator; only the f64 summation order
differs); the adjoint agrees with central differences of the device's forward (Russian
roulette off, step 1e-3) within 1e-3 relative; at full size the fit's first gradient is the
batched adjoint of its own dL/dI and the loss falls."""

import numpy as np
import pytest

from lumenwright_amd import diff, frames

pytestmark = pytest.mark.gpu

WORLD = frames.WorldSpec(8, 8, 24, cube_types=24, cutout_types=4, water_level=62,
                         emitter_fraction=0.01, cutout_fraction=0.05)


def smooth(p):
    q = p.with_values()
    q.rr_off = True
    return q


@pytest.fixture(scope="module")
def small_fit():
    return diff.config4_problem(n_views=3, size=(64, 36), spp=4, n_groups=16, world=WORLD)


@pytest.mark.parametrize("rr_off", [False, True])
def test_fit_gradient_equals_the_cpu_reference(hip_backend, oracle, small_fit, rr_off):
    bp = small_fit
    p = smooth(bp.start) if rr_off else bp.start.with_values()
    rng = np.random.default_rng(12)
    Gs = [rng.standard_normal(int(v["frame"][0]) * int(v["frame"][1]) * 3).astype("<f4")
          for v, _ in bp.views]
    n = p.n_grads
    want = np.zeros(n)
    for (v, spp), G in zip(bp.views, Gs):
        want += oracle.adjoint(frames.HqFrame(bp.base.scene, v, spp, bp.base.sky).view(), spp,
                               p.desc(), n, G)
    sc = hip_backend.upload(bp.base)
    try:
        got_batch = hip_backend.adjoint_batch(sc, bp.views, p.desc(), n, Gs)
        got_one = sum(hip_backend.adjoint(sc, v, spp, p.desc(), n, G)
                      for (v, spp), G in zip(bp.views, Gs))
        imgs = hip_backend.render_diff_batch(sc, bp.views, bp.truth.desc())
    finally:
        sc.free()
    scale = np.abs(want).max()
    print(f"fit gradient (rr_off={rr_off}): batch {np.abs(got_batch - want).max() / scale:.2e}, "
          f"per view {np.abs(got_one - want).max() / scale:.2e} of the largest component")
    assert scale > 0 and np.count_nonzero(want[3 * p.n_palette:3 * p.n_palette + p.n_intensity])
    assert np.abs(got_batch - want).max() <= 1e-6 * scale
    assert np.abs(got_one - want).max() <= 1e-6 * scale
    for (v, spp), img in zip(bp.views, imgs):  # the forward at the truth
        want_img = oracle.render_diff(frames.HqFrame(bp.base.scene, v, spp, bp.base.sky).view(),
                                      spp, bp.truth.desc()).reshape(-1, 3)
        assert np.array_equal(img.view("<u4"), want_img.view("<u4"))


def test_fit_gradient_matches_central_differences(hip_backend):
    # One view of 256x144 at 16 spp: at this size the f32 rounding of the samples moves each
    # difference by about 2e-4 of its value at step 1e-3.
    bp = diff.config4_problem(n_views=1, size=(256, 144), spp=16, n_groups=16, world=WORLD)
    (v, spp), = bp.views
    start = smooth(bp.start)
    G = np.random.default_rng(6).standard_normal(256 * 144 * 3).astype("<f4")
    sc = hip_backend.upload(bp.base)
    try:
        adj = hip_backend.adjoint(sc, v, spp, start.desc(), start.n_grads, G)

        def J(p):
            _, img64 = hip_backend.render_diff(sc, v, spp, p.desc(), f64=True)
            return float(np.dot(img64.reshape(-1), G.astype("<f8")))

        h = 1e-3
        fd = []
        for i in range(3 * start.n_palette):
            pp, pm = start.palette.ravel().copy(), start.palette.ravel().copy()
            pp[i] += h
            pm[i] -= h
            fd.append((J(smooth(start.with_values(palette=pp)))
                       - J(smooth(start.with_values(palette=pm)))) / (2 * h))
        for j in range(start.n_intensity):
            ip, im = start.intensity.copy(), start.intensity.copy()
            ip[j] += h
            im[j] -= h
            fd.append((J(smooth(start.with_values(intensity=ip)))
                       - J(smooth(start.with_values(intensity=im)))) / (2 * h))
    finally:
        sc.free()
    fd = np.array(fd)
    a = adj[:len(fd)]
    rel = np.abs(a - fd) / np.maximum(np.abs(fd), 1e-2 * np.abs(fd).max())
    k = int(np.argmax(rel))
    print(f"fit world vs differences: worst rel {rel.max():.2e} at {k} ({a[k]:.6g} vs {fd[k]:.6g}), "
          f"{int(np.count_nonzero(fd[3 * start.n_palette:]))} intensity groups seen")
    assert np.count_nonzero(fd[3 * start.n_palette:]) > 0
    assert rel.max() <= 1e-3


def test_full_size_fit_steps(hip_backend):
    # Config 4 in full: 32 views of 640x360 at 128 spp, 6 bounces, 192 albedo factors and 128
    # intensity groups. The first step's gradient is the batched adjoint of dL/dI = 2 (I - T) / N
    # at the start, and three steps lower the loss.
    bp = diff.config4_problem()
    sc = hip_backend.upload(bp.base)
    fit = None
    try:
        targets = hip_backend.render_diff_batch(sc, bp.views, bp.truth.desc())
        fit = hip_backend.fit(sc, bp.views, bp.start.desc(), targets)
        first = fit.step()
        imgs = hip_backend.render_diff_batch(sc, bp.views, bp.start.desc())
        n = sum(i.size for i in imgs)
        dldis = [(2.0 / n) * (i.astype("<f8") - t) for i, t in zip(imgs, targets)]
        want = hip_backend.adjoint_batch(sc, bp.views, bp.start.desc(), bp.start.n_grads, dldis)
        _, _, g = fit.params()
        loss0 = sum(float(np.sum((i.astype("<f8") - t) ** 2)) for i, t in zip(imgs, targets)) / n
        steps = [first] + [fit.step() for _ in range(2)]
    finally:
        if fit is not None:
            fit.free()
        sc.free()
    print("full-size fit: losses", [s["loss"] for s in steps], "forward/adjoint ms",
          [(round(s["forward_ms"], 1), round(s["adjoint_ms"], 1)) for s in steps],
          "gradient rel", np.abs(g - want).max() / np.abs(want).max())
    assert first["samples"] == 32 * 640 * 360 * 128
    assert abs(first["loss"] - loss0) <= 1e-6 * loss0
    assert np.abs(g - want).max() <= 1e-3 * np.abs(want).max()
    assert steps[1]["loss"] < steps[0]["loss"] and steps[2]["loss"] < steps[1]["loss"]
