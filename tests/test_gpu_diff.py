"""Differentiable rendering on the device (lw_hip_render_diff, lw_hip_adjoint) through the C
ABI, against the CPU reference (oracle/lw_oracle_diff.hpp) and central differences:

  * the forward at identity parameters is render_hq's image bit for bit, and at other
    parameters the CPU reference's image;
  * the adjoint equals the CPU reference gradient (the same samples and estimator; only the
    f64 summation order differs);
  * at the cfg3 size (256x144, 64 spp, 4 bounces) the adjoint agrees with central
    differences of the device's own forward, with Russian roulette off: palette and
    intensity factors (step 1e-3) and turbidity (step 3e-4) within 1e-3 relative; the sun
    angles, whose per-sample differences jump where a shadow edge crosses the sun's cone,
    agree with a difference reference averaged over 1024 seeds (standard error below 1e-3
    of its value) within three standard errors, and meet the closed form of a sunlit floor
    within 1e-3.
"""

import numpy as np
import pytest

from lumenwright_amd import diff, frames
from lumenwright_amd.abi import EMITTER_SAMPLING_ALL, NO_GROUP

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def problem():
    return diff.config3_problem()


def smooth(p):
    q = p.with_values()
    q.rr_off = True
    return q


def test_identity_forward_is_render_hq(hip_backend, problem):
    cfg = frames.config(3, width=96, height=54, spp=4)
    fr = diff.frame_at(cfg, problem.scene, problem.start.sky)
    sc = hip_backend.upload(fr)
    a = hip_backend.render_hq_scene(sc, fr.params, 4)
    b = hip_backend.render_diff(sc, fr.params, 4, problem.start.desc())
    assert np.array_equal(a.view("<u4"), b.view("<u4"))
    sc.free()


def test_identity_forward_is_render_hq_with_water_and_plants(hip_backend):
    cfg = frames.config(1, width=64, height=36, spp=2, world_chunks_x=6, world_chunks_z=6)
    fr = frames.make_frame(cfg)
    s = fr.scene
    p = diff.identity(diff.block_type_groups(s), diff.emitter_groups(s, 8),
                      int(s.quad_type[s.quad_type != NO_GROUP].max()) + 1, 8)
    sc = hip_backend.upload(fr)
    a = hip_backend.render_hq_scene(sc, fr.params, 2)
    b = hip_backend.render_diff(sc, fr.params, 2, p.desc())
    assert np.array_equal(a.view("<u4"), b.view("<u4"))
    sc.free()


def test_forward_matches_the_cpu_reference(hip_backend, oracle, problem):
    cfg = frames.config(3, width=64, height=36, spp=4)
    fr = diff.frame_at(cfg, problem.scene, problem.target.sky)
    sc = hip_backend.upload(fr)
    got = hip_backend.render_diff(sc, fr.params, 4, problem.target.desc())
    want = oracle.render_diff(fr.view(), 4, problem.target.desc()).reshape(-1, 3)
    assert np.array_equal(got.view("<u4"), want.view("<u4"))
    sc.free()


@pytest.mark.parametrize("rr_off", [False, True])
def test_adjoint_matches_the_cpu_reference(hip_backend, oracle, problem, rr_off):
    cfg = frames.config(3, width=64, height=36, spp=8)
    fr = diff.frame_at(cfg, problem.scene, problem.start.sky)
    derivs = diff.sky_derivatives(cfg, problem.scene, problem.start.sky, 3e-4)
    p = smooth(problem.start) if rr_off else problem.start.with_values()
    target = oracle.render_diff(diff.frame_at(cfg, problem.scene, problem.target.sky).view(), 8,
                                problem.target.desc()).reshape(-1, 3)
    sc = hip_backend.upload(fr)
    img = hip_backend.render_diff(sc, fr.params, 8, p.desc())
    loss, dldi = diff.l2_loss(img, target)
    got = hip_backend.adjoint(sc, fr.params, 8, p.desc(derivs), p.n_grads, dldi)
    want = oracle.adjoint(fr.view(), 8, p.desc(derivs), p.n_grads, dldi)
    scale = np.abs(want).max()
    assert scale > 0 and loss > 0
    assert np.abs(got - want).max() <= 1e-6 * scale, np.abs(got - want).max() / scale
    sc.free()


@pytest.fixture(scope="module")
def full_size(hip_backend, problem):
    """The cfg3 frame at full size, its adjoint with Russian roulette off, and a weight
    image G: sum(G * I) is the function the differences take."""
    cfg = frames.CONFIGS[3]
    fr = diff.frame_at(cfg, problem.scene, problem.start.sky)
    G = np.random.default_rng(11).standard_normal(cfg.width * cfg.height * 3).astype("<f4")
    start = smooth(problem.start)
    derivs = diff.sky_derivatives(cfg, problem.scene, problem.start.sky, 3e-4)
    sc = hip_backend.upload(fr)
    adj = hip_backend.adjoint(sc, fr.params, cfg.spp, start.desc(derivs), start.n_grads, G)
    yield cfg, fr, sc, G, start, adj
    sc.free()


def J(backend, sc, params, spp, p, G):
    # sum(G * I) over the f64 means: the f32 image's rounding would add about 4e-3 of noise to
    # each difference at this size.
    _, img64 = backend.render_diff(sc, params, spp, p.desc(), f64=True)
    return float(np.dot(img64.reshape(-1), G.astype("<f8")))


def test_adjoint_matches_central_differences_at_full_size(hip_backend, problem, full_size):
    cfg, fr, sc, G, start, adj = full_size
    h = 1e-3
    fd, idx = [], []
    for i in range(3 * start.n_palette):
        pp, pm = start.palette.ravel().copy(), start.palette.ravel().copy()
        pp[i] += h
        pm[i] -= h
        fd.append((J(hip_backend, sc, fr.params, cfg.spp, smooth(start.with_values(palette=pp)), G)
                   - J(hip_backend, sc, fr.params, cfg.spp, smooth(start.with_values(palette=pm)), G)) / (2 * h))
        idx.append(i)
    for j in range(start.n_intensity):
        ip, im = start.intensity.copy(), start.intensity.copy()
        ip[j] += h
        im[j] -= h
        fd.append((J(hip_backend, sc, fr.params, cfg.spp, smooth(start.with_values(intensity=ip)), G)
                   - J(hip_backend, sc, fr.params, cfg.spp, smooth(start.with_values(intensity=im)), G)) / (2 * h))
        idx.append(3 * start.n_palette + j)
    fd, idx = np.array(fd), np.array(idx)
    a = adj[idx]
    rel = np.abs(a - fd) / np.maximum(np.abs(fd), 1e-2 * np.abs(fd).max())
    print("palette+intensity worst rel", rel.max(), "max |fd|", np.abs(fd).max())
    assert rel.max() <= 1e-3
    # turbidity: the sky cache changes, so each side is its own upload
    ht = 3e-4
    vals = []
    for sgn in (1, -1):
        f = diff.frame_at(cfg, problem.scene, start.sky + np.array([sgn * ht, 0.0, 0.0]))
        s2 = hip_backend.upload(f)
        vals.append(J(hip_backend, s2, f.params, cfg.spp, start, G))
        s2.free()
    fd_t = (vals[0] - vals[1]) / (2 * ht)
    print("turbidity", adj[-3], fd_t)
    assert abs(adj[-3] - fd_t) <= 1e-3 * abs(fd_t)


def test_sun_angle_gradient_agrees_with_a_1e3_difference_reference(hip_backend, problem):
    # The sun's altitude and azimuth at the cfg3 size, on the image's total radiance (G = 1,
    # so that the gradient grows with the pixel count and the noise with its root). A sample's
    # central difference jumps where a shadow edge crosses its cone direction, so one seed's
    # difference is noisy (0.8% and 2.8% at h = 1e-2) while its mean over seeds, with common
    # random numbers on both sides, converges. 1024 seeds bring the reference's standard error
    # under 1e-3 of its value, and the adjoint's mean (the rim estimate, 0.5% a seed) agrees
    # with it within three standard errors of the difference.
    cfg = frames.CONFIGS[3]
    start = smooth(problem.start)
    G = np.ones(cfg.width * cfg.height * 3, "<f4")
    h = 1e-2
    n_seeds = 1024
    base = diff.frame_at(cfg, problem.scene, start.sky)
    derivs = diff.sky_derivatives(cfg, problem.scene, start.sky, 3e-4)
    sides = {}
    for k in (1, 2):
        for sgn in (1, -1):
            e = np.zeros(3)
            e[k] = sgn * h
            f = diff.frame_at(cfg, problem.scene, start.sky + e)
            sides[(k, sgn)] = (f, hip_backend.upload(f))
    sc = hip_backend.upload(base)
    adj, fd = [], []
    try:
        for seed in range(n_seeds):
            pr = base.params.copy()
            pr["misc_u"][0] = 100 + seed
            adj.append(hip_backend.adjoint(sc, pr, cfg.spp, start.desc(derivs), start.n_grads, G)[-2:])
            row = []
            for k in (1, 2):
                v = []
                for sgn in (1, -1):
                    f, s2 = sides[(k, sgn)]
                    q = f.params.copy()
                    q["misc_u"][0] = 100 + seed
                    v.append(float(hip_backend.render_diff(s2, q, cfg.spp, start.desc(), f64=True)[1].sum()))
                row.append((v[0] - v[1]) / (2 * h))
            fd.append(row)
    finally:
        for _, s2 in sides.values():
            s2.free()
        sc.free()
    adj, fd = np.array(adj), np.array(fd)
    ref, se_ref = fd.mean(axis=0), fd.std(axis=0, ddof=1) / np.sqrt(n_seeds)
    d = adj - fd
    dm, se_d = d.mean(axis=0), d.std(axis=0, ddof=1) / np.sqrt(n_seeds)
    print(f"sun altitude, azimuth: adjoint {adj.mean(axis=0)}, differences {ref} (standard error "
          f"{se_ref / np.abs(ref)} relative); adjoint - differences {dm / np.abs(ref)} relative, "
          f"{np.abs(dm) / se_d} standard errors")
    assert np.all(se_ref <= 1e-3 * np.abs(ref))
    assert np.all(np.abs(dm) <= 3 * se_d)


def test_unsupported_and_invalid_requests_fail(hip_backend, problem):
    import lumenwright_amd as lw
    cfg = frames.config(3, width=32, height=18, spp=2)
    fr = diff.frame_at(cfg, problem.scene, problem.start.sky)
    sc = hip_backend.upload(fr)
    try:
        p_all = fr.params.copy()
        p_all["emitter"][0] = EMITTER_SAMPLING_ALL
        with pytest.raises(lw.BackendError) as e:
            hip_backend.render_diff(sc, p_all, 2, problem.start.desc())
        assert e.value.code == lw.backend.LW_EUNSUPPORTED
        deep = fr.params.copy()
        deep["frame"][3] = 17
        g = np.zeros(32 * 18 * 3, "<f4")
        with pytest.raises(lw.BackendError) as e:
            hip_backend.adjoint(sc, deep, 2, problem.start.desc(), problem.start.n_grads, g)
        assert e.value.code == lw.backend.LW_EINVAL
        zero = problem.start.with_values(palette=np.zeros_like(problem.start.palette))
        with pytest.raises(lw.BackendError) as e:
            hip_backend.render_diff(sc, fr.params, 2, zero.desc())
        assert e.value.code == lw.backend.LW_EINVAL
        bad = problem.start.with_values()
        bad.quad_palette = bad.quad_palette.copy()
        bad.quad_palette[0] = 1000
        with pytest.raises(lw.BackendError):
            hip_backend.render_diff(sc, fr.params, 2, bad.desc())
    finally:
        sc.free()


@pytest.fixture(scope="module")
def batch_problem():
    return diff.config4_problem(n_views=4, size=(96, 54), spp=8, n_groups=16,
                                world=frames.WorldSpec(8, 8, 24, cube_types=24, cutout_types=4,
                                                       water_level=62, emitter_fraction=0.01,
                                                       cutout_fraction=0.05))


@pytest.fixture(scope="module")
def batch_problem_64():
    # 64 spp per view in one pass: the passes run a wave per pixel's 64 samples
    return diff.config4_problem(n_views=3, size=(72, 40), spp=64, n_groups=16,
                                world=frames.WorldSpec(8, 8, 24, cube_types=24, cutout_types=4,
                                                       water_level=62, emitter_fraction=0.01,
                                                       cutout_fraction=0.05))


@pytest.fixture(params=["8spp", "64spp"])
def any_batch_problem(request, batch_problem, batch_problem_64):
    return batch_problem if request.param == "8spp" else batch_problem_64


def test_batched_forward_equals_single_views(hip_backend, any_batch_problem):
    bp = any_batch_problem
    sc = hip_backend.upload(bp.base)
    outs = hip_backend.render_diff_batch(sc, bp.views, bp.truth.desc())
    for (p, spp), img in zip(bp.views, outs):
        one = hip_backend.render_diff(sc, p, spp, bp.truth.desc())
        assert np.array_equal(img.view("<u4"), one.view("<u4"))
    sc.free()


def test_batched_adjoint_is_the_sum_of_the_views(hip_backend, any_batch_problem):
    bp = any_batch_problem
    sc = hip_backend.upload(bp.base)
    rng = np.random.default_rng(2)
    Gs = [rng.standard_normal(int(p["frame"][0]) * int(p["frame"][1]) * 3).astype("<f4")
          for p, _ in bp.views]
    n = bp.start.n_grads
    got = hip_backend.adjoint_batch(sc, bp.views, bp.truth.desc(), n, Gs)
    want = sum(hip_backend.adjoint(sc, p, spp, bp.truth.desc(), n, G)
               for (p, spp), G in zip(bp.views, Gs))
    assert np.abs(want).max() > 0
    assert np.abs(got - want).max() <= 1e-9 * np.abs(want).max()
    sc.free()


def test_fit_reduces_the_loss_and_moves_towards_the_truth(hip_backend, batch_problem):
    bp = batch_problem
    sc = hip_backend.upload(bp.base)
    targets = hip_backend.render_diff_batch(sc, bp.views, bp.truth.desc())
    fit = hip_backend.fit(sc, bp.views, bp.start.desc(), targets, lr=2e-2)
    first = fit.step()
    # The first step's gradient is the batched adjoint of dL/dI = 2 (I - T) / N at the start.
    imgs = hip_backend.render_diff_batch(sc, bp.views, bp.start.desc())
    n = sum(i.size for i in imgs)
    dldis = [(2.0 / n) * (i.astype("<f8") - t) for i, t in zip(imgs, targets)]
    want = hip_backend.adjoint_batch(sc, bp.views, bp.start.desc(), bp.start.n_grads, dldis)
    _, _, g = fit.params()
    assert np.abs(g - want).max() <= 1e-3 * np.abs(want).max()
    loss0 = first["loss"]
    assert abs(loss0 - sum(float(np.sum((i.astype("<f8") - t) ** 2)) for i, t in zip(imgs, targets)) / n) <= 1e-6 * loss0
    for _ in range(40):
        last = fit.step()
    pal, inten, _ = fit.params()
    fitted = bp.start.with_values(pal, inten)
    e0 = diff.parameter_error(bp.start, bp.truth)
    e1 = diff.parameter_error(fitted, bp.truth)
    print("loss", loss0, "->", last["loss"], "error", e0, "->", e1)
    assert last["loss"] < 0.5 * loss0
    assert e1[0] < e0[0]
    fit.free()
    sc.free()


@pytest.mark.parametrize("alt,az", [(0.9, 0.4), (1.3, 2.5)])
def test_sun_angle_gradient_meets_the_closed_form_to_1e3(hip_backend, alt, az):
    # A white floor under the sun alone (no occlusion, the cone wholly above the horizon): the
    # image is (1 + cos a) / 2 * sin(altitude), so d/d altitude of the image's sum is
    # pixels * 3 * (1 + cos a) / 2 * cos(altitude) and d/d azimuth is 0. The device's rim
    # estimate over 32 x 32 pixels x 16384 samples (its relative spread is about 0.5% at 64
    # samples a pixel, so about 3e-4 here) meets both within 1e-3 of the altitude term.
    from test_diff_oracle import floor_desc
    import scenes
    radius = 0.05
    s, dsun = scenes.floor_scene(alt, az, radius)
    w, h = int(s.params["frame"][0]), int(s.params["frame"][1])
    spp = 16384
    p, desc = floor_desc(dsun)
    sc = hip_backend.upload(s)
    try:
        g = hip_backend.adjoint(sc, s.params, spp, desc, p.n_grads, np.ones(w * h * 3, "<f4"))
        img = hip_backend.render_hq_scene(sc, s.params, 256)
    finally:
        sc.free()
    cos_a = float(np.float32(np.cos(radius)))
    want_img = (1 + cos_a) / 2 * np.sin(alt)
    want = w * h * 3 * (1 + cos_a) / 2 * np.cos(alt)
    print(f"alt {alt}: image {img.mean():.6f} vs {want_img:.6f}; d/d altitude {g[-2]:.2f} vs "
          f"{want:.2f} ({(g[-2] - want) / want:.2e}); d/d azimuth {g[-1]:.3f}")
    assert abs(img.mean() - want_img) <= 2e-3 * want_img
    assert abs(g[-2] - want) <= 1e-3 * abs(want)
    assert abs(g[-1]) <= 1e-3 * abs(want)
    assert g[-3] == 0.0
