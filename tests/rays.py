"""Seeded ray sets for traversal parity: camera rays, rays from inside and outside the grid
(axis-aligned and grazing ones among them), secondary rays leaving surfaces, with mixed
tmin/tmax -- the shapes of query render_sample makes (integrator.rs:572,684,402)."""

import numpy as np


def _unit(v):
    v = np.asarray(v, dtype=np.float64)
    return (v / np.linalg.norm(v, axis=-1, keepdims=True)).astype(np.float32)


def camera_rays(params, n, rng):
    w, h = int(params["frame"][0]), int(params["frame"][1])
    px = rng.integers(0, w, n) + rng.random(n)
    py = rng.integers(0, h, n) + rng.random(n)
    th, asp = float(params["cam_pos"][3]), float(params["cam_fwd"][3])
    sx = (2.0 * px / w - 1.0) * asp * th
    sy = (1.0 - 2.0 * py / h) * th
    f, r, u = (np.array(params[k][:3], dtype=np.float64) for k in ("cam_fwd", "cam_right", "cam_up"))
    d = f[None] + r[None] * sx[:, None] + u[None] * sy[:, None]
    o = np.repeat(np.array(params["cam_pos"][:3], dtype=np.float32)[None], n, 0)
    return o, _unit(d)


def interior_rays(params, n, rng, y_lo=30.0, y_hi=110.0):
    dims = params["grid_dims"][:3].astype(np.float64) * 16
    o = np.stack([rng.random(n) * dims[0], y_lo + rng.random(n) * (y_hi - y_lo),
                  rng.random(n) * dims[2]], 1).astype(np.float32)
    d = rng.normal(size=(n, 3))
    # axis-aligned and grazing directions: zero components, tiny components
    k = n // 10
    d[:k, rng.integers(0, 3)] = 0.0
    d[k:2 * k, 1] = 1e-7 * rng.normal(size=k)
    axis = rng.integers(0, 3, k)
    d[2 * k:3 * k] = 0.0
    d[2 * k + np.arange(k), axis] = rng.choice([-1.0, 1.0], k)
    return o, _unit(d)


def outside_rays(params, n, rng):
    dims = params["grid_dims"][:3].astype(np.float64) * 16
    c = dims / 2
    dirs = _unit(rng.normal(size=(n, 3)))
    o = (c[None] - dirs.astype(np.float64) * (np.linalg.norm(dims) * 0.8)).astype(np.float32)
    target = c[None] + (rng.random((n, 3)) - 0.5) * dims[None] * 0.6
    return o, _unit(target - o)


def secondary_rays(o, d, hits, quads, rng, eps=1e-3):
    """Rays leaving the hit points of (o, d) into the hemisphere of the hit quad's normal."""
    m = hits["hit"] == 1
    o, d, hits = o[m], d[m], hits[m]
    q = quads[hits["quad"]]
    nrm = np.cross(q["ev"][:, :3].astype(np.float64), q["eu"][:, :3].astype(np.float64))
    nrm = nrm / np.linalg.norm(nrm, axis=1, keepdims=True)
    flip = np.sum(nrm * d, 1) > 0
    nrm[flip] *= -1
    pos = o.astype(np.float64) + d.astype(np.float64) * hits["t"][:, None].astype(np.float64)
    so = (pos + nrm * eps).astype(np.float32)
    sd = rng.normal(size=so.shape)
    sd += nrm * 1.5
    return so, _unit(sd)


def ray_mix(params, quads, trace_fn, n, seed):
    """All of the above, with tmin in {0, 1e-3, U[0, 4)} and tmax in {inf, U[1, 200)}."""
    rng = np.random.default_rng(seed)
    parts = [camera_rays(params, n, rng), interior_rays(params, n, rng), outside_rays(params, n, rng)]
    o0, d0 = parts[0]
    h0 = trace_fn(o0, d0, np.zeros(len(o0), np.float32), np.full(len(o0), np.inf, np.float32))
    parts.append(secondary_rays(o0, d0, h0, quads, rng))
    o = np.concatenate([p[0] for p in parts])
    d = np.concatenate([p[1] for p in parts])
    m = len(o)
    tmin = np.where(rng.random(m) < 0.5, 0.0, np.where(rng.random(m) < 0.5, 1e-3,
                                                       rng.random(m) * 4)).astype(np.float32)
    tmax = np.where(rng.random(m) < 0.7, np.inf, 1.0 + rng.random(m) * 199).astype(np.float32)
    return o, d, tmin, tmax
