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
    agree in the mean over seeds within their standard errors.
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


def test_sun_angle_gradient_agrees_with_differences_in_the_mean(hip_backend, problem):
    # Per seed, central differences of the sun angles jump wherever a shadow edge crosses a
    # sampled cone direction, and the rim estimate is noisy; their means over seeds must
    # agree within the combined standard error.
    cfg = frames.config(3, width=128, height=72, spp=64)
    start = smooth(problem.start)
    G = np.random.default_rng(5).standard_normal(cfg.width * cfg.height * 3).astype("<f4")
    h = 1e-3
    frames_pm = {}
    for k in (1, 2):
        for sgn in (1, -1):
            e = np.zeros(3)
            e[k] = sgn * h
            frames_pm[(k, sgn)] = diff.frame_at(cfg, problem.scene, start.sky + e)
    base = diff.frame_at(cfg, problem.scene, start.sky)
    derivs = diff.sky_derivatives(cfg, problem.scene, start.sky, 3e-4)
    scenes_pm = {key: hip_backend.upload(f) for key, f in frames_pm.items()}
    sc = hip_backend.upload(base)
    adj, fd = [], []
    for seed in range(8):
        pr = base.params.copy()
        pr["misc_u"][0] = 100 + seed
        adj.append(hip_backend.adjoint(sc, pr, cfg.spp, start.desc(derivs), start.n_grads, G)[-2:])
        row = []
        for k in (1, 2):
            jp = frames_pm[(k, 1)].params.copy()
            jm = frames_pm[(k, -1)].params.copy()
            jp["misc_u"][0] = jm["misc_u"][0] = 100 + seed
            row.append((J(hip_backend, scenes_pm[(k, 1)], jp, cfg.spp, start, G)
                        - J(hip_backend, scenes_pm[(k, -1)], jm, cfg.spp, start, G)) / (2 * h))
        fd.append(row)
    for s in list(scenes_pm.values()) + [sc]:
        s.free()
    adj, fd = np.array(adj), np.array(fd)
    n = len(adj)
    se = np.sqrt(adj.var(axis=0, ddof=1) / n + fd.var(axis=0, ddof=1) / n)
    diffm = np.abs(adj.mean(axis=0) - fd.mean(axis=0))
    print("sun adjoint", adj.mean(axis=0), "differences", fd.mean(axis=0), "se", se)
    assert np.all(diffm <= 4 * se + 1e-3 * np.abs(fd.mean(axis=0)))


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


def test_batched_forward_equals_single_views(hip_backend, batch_problem):
    bp = batch_problem
    sc = hip_backend.upload(bp.base)
    outs = hip_backend.render_diff_batch(sc, bp.views, bp.truth.desc())
    for (p, spp), img in zip(bp.views, outs):
        one = hip_backend.render_diff(sc, p, spp, bp.truth.desc())
        assert np.array_equal(img.view("<u4"), one.view("<u4"))
    sc.free()


def test_batched_adjoint_is_the_sum_of_the_views(hip_backend, batch_problem):
    bp = batch_problem
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
