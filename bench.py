#!/usr/bin/env python3
"""Benchmark: Msamples/s of the HQ path tracer on one MI355X.

Default (config 1): a step is one render_hq of 1280 x 720 at 256 spp, 8 bounces, over the
24 x 24-chunk world with water and emitters under the game dusk sky, with the scene
resident in HBM. W untimed warmup steps, then K timed steps; one JSON line.

--config 2: the batched workload -- 64 seeded views of the config-1 world with sizes from
{160x90 .. 1280x720} and spp from {16 .. 1024} -- rendered once as one batch
(lw_hip_render_hq_batch) and once as 64 plain render_hq calls; a step renders all 64 views
both ways. The images are identical (same samples, same summation order), so the two are
compared at equal image error. The views are also rendered once through the reference's
per-view launch loop (render/cuda.rs:132-169) on the device (LW_FRAME_FORM=launches), again
with identical images.

--config 3: the differentiable render of the 8 x 8-chunk world with 16 emitter blocks
(256 x 144, 64 spp, 4 bounces): a step is one forward and one adjoint (path replay) of an L2
image loss over 115 parameters; the line carries the adjoint / forward kernel-time ratio.

--config 4: the batched inverse fit -- 32 views of the config-1 world at 640 x 360, 128 spp,
6 bounces, 192 albedo factors and 128 emitter-intensity groups -- a step is one fit step on
the device (batched forward, loss, batched adjoint, Adam update); --fit-steps runs a longer
fit and reports the parameter error against the ground truth.

Besides the throughput the config-1 line carries:
  roofline      the path-tracing kernels' algorithmic scene bytes per launch (per-sample
                bytes counted once by the instrumented oracle, profiles/b_alg.json, times
                the samples of a launch) over their mean launch time (HIP events on the
                library's stream), against the 8 TB/s HBM peak, with the HBM bytes per launch
                the kernel's counters measured (profiles/traffic.json);
  cpu_baseline  the oracle (the C++ restatement of the reference's CPU backend) on the
                host's cores over a bounded slice of the same frame (the first S samples of
                every pixel), timed in this run;
  mrays_per_s   the trace calls of a step (camera, extension, shadow and emitter rays),
                counted on the device by an untimed replay of the step's paths
                (lw_hip_count_traces), over the step's time; mrays_per_s_oracle_rate is
                Msamples/s times the oracle's trace calls per sample, as a cross-check;
  post          the device tonemap and tone mix of the last frame (kernel time, GB/s);
  scene_build   the world's layout made on the device from its voxel volume
                (lw_hip_scene_build), and the oracle's CPU restatement of it;
  sky_prep      the game's sky maps, Chunky's sky cache and their blend on the device, and
                the oracle's CPU game_sky_maps;
  parity        the device's image of the frame at the CPU baseline's last spp against the
                image that pass made (pixels bit for bit, MAE), and the MAE of both against
                the f64 render on a subset of rows;
and the metric's other two parts (leave them out with --no-metric-parts):
  batched       config 2 in full: 64 views as one batch and as 64 plain calls, the ratio
                batched_over_plain and whether the images are identical; and against the
                reference's per-view launch loop, batched_over_launch_loop;
  adjoint       config 3's adjoint / forward kernel-time ratio, and one config-4 fit step's
                forward and adjoint times.
"""

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
METRIC = "Msamples/s at 720p 8-bounce hq; batched vs plain; adjoint/forward time ratio"


def log(msg):
    print("# " + msg, file=sys.stderr, flush=True)


def host_cores() -> int:
    n = len(os.sched_getaffinity(0))
    cap = os.environ.get("OMP_NUM_THREADS")
    if cap and cap.isdigit():
        n = min(n, int(cap))
    return max(1, n)


def cpu_baseline(frame, seconds: float):
    """The oracle's CpuBackend::render_hq over the frame's first spp samples, doubling spp
    until the total time reaches `seconds`. Returns the figure, the last pass's image and its
    spp (the parity check renders the same on the device)."""
    from oracle import ffi
    ffi.build()
    from lumenwright_amd.abi import HqFrame
    cores = host_cores()
    v = HqFrame(frame.scene, frame.params, 1, frame.sky).view()
    w, h = int(frame.params["frame"][0]), int(frame.params["frame"][1])
    spp, total_s, total_n = 1, 0.0, 0
    while True:
        t = time.perf_counter()
        img = ffi.render_hq(v, spp, threads=cores)
        dt = time.perf_counter() - t
        total_s += dt
        total_n += w * h * spp
        if total_s >= seconds or spp >= frame.spp:
            break
        spp = min(spp * 2, frame.spp)
    return {"value": total_n / total_s / 1e6, "unit": "Msamples/s", "cores": cores,
            "kind": "port",
            "sample": f"{w}x{h}, passes of the first 1, 2, 4, ... {spp} samples of every "
                      f"pixel: {total_n} samples in {total_s:.1f} s"}, img, spp


def parity_block(b, scene, frame, cpu_img, spp: int, n_rows: int = 16) -> dict:
    """The headline frame against the CPU backend: the device's image at the CPU baseline's
    last spp, compared pixel by pixel with the oracle image that pass made (untimed), and the
    error of both against the f64 render (oracle/lw_oracle_f64.cpp) on n_rows evenly spaced
    rows of the frame."""
    from oracle import ffi
    from lumenwright_amd.abi import HqFrame
    w, h = int(frame.params["frame"][0]), int(frame.params["frame"][1])
    dev = b.render_hq_scene(scene, frame.params, spp).reshape(h, w, 3)
    same = np.all(dev.view("<u4") == cpu_img.view("<u4"), axis=2)
    rows = np.linspace(0, h - 1, n_rows).round().astype(int)
    t = time.perf_counter()
    f64 = ffi.render_rows_f64(HqFrame(frame.scene, frame.params, spp, frame.sky).view(), spp, rows,
                              threads=host_cores())
    f64_s = time.perf_counter() - t
    err = lambda img: float(np.abs(img[rows].astype(np.float64) - f64).mean())
    return {"spp": spp, "pixels": w * h, "bit_exact_pixels": int(same.sum()),
            "mae_vs_cpu_backend": float(np.abs(dev.astype(np.float64) - cpu_img).mean()),
            "mae_vs_f64": {"device": err(dev), "cpu_backend": err(cpu_img),
                           "f64_mean": float(f64.mean()), "rows": [int(r) for r in rows],
                           "f64_seconds": f64_s}}


def roofline(config, samples, trace_ms, launches, steps=1):
    """The dominant kernel against HBM: achieved = algorithmic scene bytes per launch (the
    instrumented oracle's bytes per sample times the launch's samples) over the mean launch
    time; traffic = the HBM bytes per launch the FETCH_SIZE / WRITE_SIZE counters measured for
    the same kernel and workload (profiles/traffic.json), when this run is that kernel."""
    b_alg_path = ROOT / "profiles" / "b_alg.json"
    if not b_alg_path.exists():
        return None
    b_alg = json.loads(b_alg_path.read_text()).get(f"config{config}")
    if not b_alg or launches == 0:
        return None
    per_launch = samples / launches
    mean_s = trace_ms / 1e3 / launches
    achieved = b_alg["bytes_per_sample"] * per_launch / mean_s / 1e9
    # one launch per step: the persistent tile kernel (the default form, not the launch ramp)
    persistent = (launches == steps and os.environ.get("LW_KERNEL", "direct") == "direct"
                  and os.environ.get("LW_FRAME_FORM") != "launches")
    traffic = None
    traffic_path = ROOT / "profiles" / "traffic.json"
    if traffic_path.exists() and persistent:
        t = json.loads(traffic_path.read_text()).get(f"config{config}")
        waves = os.environ.get("LW_TILES_WAVES", os.environ.get("LW_WAVES", "5"))
        if t and str(t.get("tiles_waves")) == waves:
            traffic = t["hbm_bytes_per_sample"] * per_launch
    return {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
            "kernel": ("k_render_tiles (one persistent launch per frame)" if persistent else
                       "path-tracing kernels of a launch (k_render | k_paths | k_wf_*)"),
            "mean_launch_ms": trace_ms / launches, "bytes_per_sample": b_alg["bytes_per_sample"],
            "note": "VALU-issue bound: the VALU issues on every cycle at 16.4% lane "
                    "utilisation, L2 hit 97.2%; HBM traffic 298 B/sample, mostly register spill "
                    "(profiles/k_render_tiles_cfg1_720p_*.md)" if config == 1 else None}


def mrays(config, msamples):
    """Mrays/s: Msamples/s times the trace calls per sample (camera, extension, shadow and
    emitter rays) the instrumented oracle counts for the workload (profiles/b_alg.json)."""
    b_alg_path = ROOT / "profiles" / "b_alg.json"
    if not b_alg_path.exists():
        return None
    b = json.loads(b_alg_path.read_text()).get(f"config{config}")
    return msamples * b["traces_per_sample"] if b else None


def kernel_form():
    form = os.environ.get("LW_KERNEL", "direct")
    w = os.environ.get("LW_WAVES")
    form += (f", persistent tiles at {os.environ.get('LW_TILES_WAVES', w or '5')} waves per SIMD, "
             f"launch ramp at {w or '6'}")
    if os.environ.get("LW_XCD_REMAP", "0") != "0":
        form += ", xcd remap"
    return form


def post_timing(b, frames, img, reps: int = 5) -> dict:
    """The device tonemap and tone mix of the rendered frame (post.rs:148-215): the best
    kernel time of a few calls (HIP events around the one kernel) and its 16 B per pixel
    (12 B of linear RGB read, 4 B of RGBA8 written) over that time."""
    n = img.shape[0]
    out = {"pixels": n}
    for name, call in (("tonemap", lambda: b.tonemap(img, 1.0, frames.TONEMAP_FILMIC_MID_GREY)),
                       ("tonemap_mix", lambda: b.tonemap_mix(img, 1.0, frames.TONEMAP_GAMMA,
                                                             frames.TONEMAP_FILMIC_MID_GREY, 0.5))):
        best = min((call(), b.last_timing()["trace_ms"])[1] for _ in range(reps))
        out[name] = {"kernel_ms": best, "GBps": 16 * n / best / 1e6}
    return out


def build_timing(b, scene, sky, cpu: bool, reps: int = 3) -> dict:
    """lw_hip_scene_build of the frame's world from its interned voxel volume: the best of a
    few builds -- the device span (HIP events from the first layout kernel to the last) and the
    whole call with the volume's upload -- and the volume's bytes over the device span. With
    `cpu`, the oracle's restatement of the same layout, timed once: single-threaded, as the
    reference runs this part (voxel.rs:305-491 and emitter_list have no rayon)."""
    vol = scene.volume
    best = None
    for _ in range(reps):
        sc, _, st = b.build_scene(vol, sky)
        sc.free()
        if best is None or st["kernel_ms"] < best["kernel_ms"]:
            best = st
    out = {"bricks": int(np.prod(vol.dims)), "stored_bricks": best["bricks"],
           "uniform_bricks": best["uniform_bricks"], "buried_bricks": best["buried_bricks"],
           "emitters": best["emitters"], "device_ms": best["kernel_ms"], "call_ms": best["wall_ms"],
           "voxel_GBps": vol.voxels.nbytes / best["kernel_ms"] / 1e6}
    if cpu:
        from oracle import ffi
        ffi.build()
        t = time.perf_counter()
        ffi.build_layout(vol)
        out["cpu_port_ms"] = (time.perf_counter() - t) * 1e3
        out["cpu_port_cores"] = 1
    return out


def sky_timing(b, frames, cpu: bool) -> dict:
    """Sky preparation on the device: the game's two maps (lw_hip_game_sky_maps, the dusk sky
    of the reference's own test, sky.rs:1462-1491), Chunky's sky cache and its blend into
    the maps; the best device time of a few calls each. With `cpu`, the oracle's
    game_sky_maps with its texels on the host's cores (sky.rs:869-872 maps them with rayon),
    timed once."""
    cam = frames.CameraRaysC()
    cam.forward[:] = [-0.83153975, -0.3506493, 0.43079773]
    cam.right[:] = [-0.45915133, 0.0, -0.888358]
    cam.up[:] = [0.0, 1.0, 0.0]
    cam.tan_half_fov, cam.aspect = 0.577, 16.0 / 9.0
    sun = frames.celestial_direction(90.29)

    def desc(bodies):
        return frames.sky_description(
            sky=(0.18, 0.24, 0.37), fog=(0.62, 0.40, 0.33), sky_fog_end=192.0,
            render_fog=(172.8, 192.0), env_fog=(0.0, 1024.0), sky_disc=True,
            sunrise=(0.98, 0.57, 0.25, 0.62), sun_angle=90.29,
            sun=(sun, (0.0, 0.0, 2.0, 2.0)) if bodies else None, star_angle=90.29,
            star_brightness=0.38 if bodies else 0.0)
    lens, light = frames.quick_sky_new(cam, 16, 9, desc(True)), frames.quick_sky_new(cam, 16, 9, desc(False))
    # the sun's 2 x 2 white sprite; one more row and column, since a texel coordinate can round
    # up to the rect's far edge
    atlas = np.full(9, 0xFFFFFFFF, "<u4")
    colour = np.array([0x92, 0x92, 0xFF], "<f4") / np.float32(255.0)
    stars = frames.star_table()
    maps, info = b.game_sky_maps(lens, light, 0.378182, colour, atlas, 2, stars)  # makes the table
    game_ms = min(b.game_sky_maps(lens, light, 0.378182, colour, atlas, 2, stars)[1]["kernel_ms"]
                  for _ in range(3))
    frame = frames.make_frame(frames.config(0, width=8, height=8, spp=1))
    p = frame.params.copy()
    p["sky_map"] = info["sizes"]
    cache_ms = min((b.sky_cache(p), b.last_timing()["trace_ms"])[1] for _ in range(3))
    blend_ms = min((b.blend_chunky_sky(p, 0.5, maps), b.last_timing()["trace_ms"])[1] for _ in range(3))
    out = {"texels": int(maps.size // 3), "game_sky_maps_ms": game_ms, "pow_table_setup_ms": info["setup_ms"],
           "sky_cache_ms": cache_ms, "blend_ms": blend_ms}
    if cpu:
        from oracle import ffi
        ffi.build()
        cores = host_cores()
        t = time.perf_counter()
        ffi.game_sky_maps(lens, light, 0.378182, colour, atlas, 2, threads=cores)
        out["cpu_port_game_sky_maps_ms"] = (time.perf_counter() - t) * 1e3
        out["cpu_port_cores"] = cores
    return out


def quick_timing(b, frames, cfg, cpu: bool, reps: int = 5) -> dict:
    """Quick mode on the device (lw_hip_render_quick) at the configuration's size: the synthetic
    world's mesh at draw distance 128; the best of a few calls' device time (inputs resident)
    and whole-call time (uploads and download included). With `cpu`, the oracle's rasterizer
    on the same frame with its tiles on the host's cores (as raster.rs runs them on rayon's
    pool), timed once, and its pixels against the device's."""
    frame, _ = frames.quick_frame(cfg)
    b.render_quick(frame)  # makes the log2 table
    dev, call = [], []
    for _ in range(reps):
        t = time.perf_counter()
        b.render_quick(frame)
        call.append((time.perf_counter() - t) * 1e3)
        dev.append(b.last_timing()["trace_ms"])
    out = {"width": frame.width, "height": frame.height, "triangles": frame.mesh.triangle_count(),
           "device_ms": min(dev), "call_ms": min(call)}
    if cpu:
        from oracle import ffi
        ffi.build()
        cores = host_cores()
        got = b.render_quick(frame)
        t = time.perf_counter()
        want = ffi.rasterize(frame, cores)
        out["cpu_port_ms"] = (time.perf_counter() - t) * 1e3
        out["cpu_port_cores"] = cores
        out["pixels_identical"] = bool(np.array_equal(got, want))
    return out


def bench_plain(args, lw, frames):
    cfg = frames.CONFIGS[args.config]
    if args.spp:
        cfg = frames.config(args.config, spp=args.spp)
    t0 = time.perf_counter()
    world = frames.build_world(cfg.world, volume=True)
    frame = frames.make_frame(cfg, world)
    log(f"frame built in {time.perf_counter() - t0:.1f} s: {cfg.name}")
    b = lw.HipBackend.open(0)
    scene = b.upload(frame)
    w, h, spp = cfg.width, cfg.height, cfg.spp
    form = kernel_form()
    tuned = None
    if args.autotune:
        tuned = b.autotune(scene, frame.params, probe_spp=4)
        bt = tuned["best"]
        form = f"{bt['form']}, waves {bt['waves']}" + (
            f", {bt['per_lane']} samples per lane" if bt["per_lane"] > 1 else "")
        if "best_tiles" in tuned:
            form += f"; persistent tiles at {tuned['best_tiles']['waves']} waves per SIMD"
        log(f"autotune: {tuned}")
    for i in range(args.warmup):
        b.render_hq_scene(scene, frame.params, spp)
        log(f"warmup {i + 1}/{args.warmup}: {b.last_timing()}")
    trace_ms, launches, acc_ms = 0.0, 0, 0.0
    t = time.perf_counter()
    for i in range(args.steps):
        img = b.render_hq_scene(scene, frame.params, spp)
        tm = b.last_timing()
        trace_ms += tm["trace_ms"]
        acc_ms += tm["accumulate_ms"]
        launches += tm["launches"]
        log(f"step {i + 1}/{args.steps}: {tm}")
    elapsed = time.perf_counter() - t
    samples = w * h * spp * args.steps
    value = samples / elapsed / 1e6
    assert np.isfinite(img).all()
    traces = b.count_traces(scene, frame.params, spp)  # untimed replay of a step's paths
    extra = {"post": post_timing(b, frames, img),
             "scene_build": build_timing(b, world, frame.sky, not args.no_cpu_baseline),
             "sky_prep": sky_timing(b, frames, not args.no_cpu_baseline),
             "quick": quick_timing(b, frames, cfg, not args.no_cpu_baseline)}
    log(f"post-processing: {extra['post']}")
    log(f"sky preparation: {extra['sky_prep']}")
    log(f"scene build: {extra['scene_build']}")
    log(f"quick mode: {extra['quick']}")
    if args.copy_probe:
        ms = b.copy_probe(args.copy_probe)
        extra["copy_probe"] = {"bytes": args.copy_probe, "ms": ms,
                               "GBps": 2 * args.copy_probe / ms / 1e6}
        log(f"copy probe: {extra['copy_probe']}")
    cpu = None
    if not args.no_cpu_baseline:
        cpu, cpu_img, cpu_spp = cpu_baseline(frame, args.cpu_seconds)
        log(f"cpu baseline: {cpu}")
        extra["parity"] = parity_block(b, scene, frame, cpu_img, cpu_spp)
        log(f"parity: {extra['parity']}")
    scene.free()
    b.close()
    if args.config == 1 and not args.no_metric_parts:
        extra["batched"] = batch_block(lw, frames, steps=1)
        log(f"batched: {extra['batched']}")
        extra["adjoint"] = {"cfg3": diff_block(lw, frames, steps=3, warmup=1),
                            "cfg4_fit_step": fit_block(lw, frames, steps=1, warmup=1)}
        log(f"adjoint: {extra['adjoint']}")
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "Msamples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 path arithmetic, f64 accumulation",
        "data": "synthetic (seeded value-noise voxel world, procedural sky)",
        "config": {"workload": cfg.name, "width": w, "height": h, "spp": spp,
                   "max_bounces": cfg.max_bounces, "kernel": form,
                   "parallelism": "1 GPU"},
        "mrays_per_s": traces * args.steps / elapsed / 1e6,
        "traces_per_sample": traces / (w * h * spp),
        "mrays_per_s_oracle_rate": mrays(args.config, value),
        "autotune": tuned,
        "kernel_ms_per_step": trace_ms / args.steps,
        "accumulate_ms_per_step": acc_ms / args.steps,
        "launches_per_step": launches / args.steps,
        "roofline": roofline(args.config, samples, trace_ms, launches, args.steps),
        "cpu_baseline": cpu,
        **extra,
    }
    if cpu:
        line["speedup_vs_cpu_baseline"] = value / cpu["value"]
    if "batched" in extra:
        line["batched_over_plain"] = extra["batched"]["batched_over_plain"]
        line["batched_over_launch_loop"] = extra["batched"]["batched_over_launch_loop"]
        line["adjoint_over_forward"] = extra["adjoint"]["cfg3"]["adjoint_over_forward"]
    return line


def batch_block(lw, frames, steps: int = 1, spp_scale: float = 1.0, bound_subset: bool = True) -> dict:
    """cfg2: the 64 seeded views of the config-1 world rendered as one batch
    (lw_hip_render_hq_batch) and as 64 plain render_hq calls, after one untimed batch; the
    wall time of each, and whether the images are identical (same samples, same summation
    order, so the two are compared at equal image error). With bound_subset, the views whose
    plain call is one persistent launch are also rendered as a batch of their own, which
    bounds what batching can gain on this view mix (full_gpu_views)."""
    cfg = frames.CONFIGS[1]
    t0 = time.perf_counter()
    scene_buf = frames.build_world(cfg.world)
    base, views = frames.batch_views(cfg, scene_buf, n_views=64, seed=3)
    if spp_scale != 1.0:
        views = [(p, max(1, int(s * spp_scale))) for p, s in views]
    samples = sum(int(p["frame"][0]) * int(p["frame"][1]) * s for p, s in views)
    log(f"64 views built in {time.perf_counter() - t0:.1f} s: {samples / 1e9:.2f} G samples")
    b = lw.HipBackend.open(0)
    scene = b.upload(base)
    try:
        batch_imgs = b.render_hq_batch(scene, views)  # warmup
        t = time.perf_counter()
        for _ in range(steps):
            batch_imgs = b.render_hq_batch(scene, views)
        t_batch = (time.perf_counter() - t) / steps
        tb = b.last_timing()
        log(f"batched: {t_batch:.2f} s per step, {tb}")
        per_view = np.zeros(len(views))
        one_launch = np.zeros(len(views), dtype=bool)
        t = time.perf_counter()
        for _ in range(steps):
            plain_imgs = []
            for i, (p, s) in enumerate(views):
                tv = time.perf_counter()
                plain_imgs.append(b.render_hq_scene(scene, p, s))
                per_view[i] += time.perf_counter() - tv
                one_launch[i] = b.last_timing()["launches"] == 1
        t_plain = (time.perf_counter() - t) / steps
        per_view /= steps
        log(f"plain: {t_plain:.2f} s per step")
        bound = None
        if bound_subset and one_launch.any():
            # The views whose plain call is already one persistent launch fill the GPU by
            # themselves; as a batch of their own they show what batching can still gain on
            # them. A batch of all 64 views does at least their work, so
            # t_plain / (their batch's time) bounds batched_over_plain from above, as if every
            # other view cost nothing.
            sub = [v for v, big in zip(views, one_launch) if big]
            b.render_hq_batch(scene, sub)  # warmup
            t = time.perf_counter()
            b.render_hq_batch(scene, sub)
            t_sub = time.perf_counter() - t
            sub_samples = sum(int(p["frame"][0]) * int(p["frame"][1]) * s for p, s in sub)
            bound = {"views": len(sub), "samples": sub_samples,
                     "plain_s": float(per_view[one_launch].sum()), "batched_s": t_sub,
                     "gain_on_these_views": float(per_view[one_launch].sum()) / t_sub,
                     "other_views_plain_s": float(per_view[~one_launch].sum()),
                     "batched_over_plain_bound": t_plain / t_sub}
            log(f"bound: {bound}")
    finally:
        scene.free()
        b.close()
    # The reference's per-view loop (render/cuda.rs:132-169: launches of 1, 2, 4, ... samples
    # per pixel, a host synchronisation and an accumulation after each) on the device: the
    # drop-in with LW_FRAME_FORM=launches, the same samples and the same images.
    saved = os.environ.get("LW_FRAME_FORM")
    os.environ["LW_FRAME_FORM"] = "launches"
    try:
        bl = lw.HipBackend.open(0)
    finally:
        if saved is None:
            os.environ.pop("LW_FRAME_FORM", None)
        else:
            os.environ["LW_FRAME_FORM"] = saved
    scene_l = bl.upload(base)
    try:
        bl.render_hq_scene(scene_l, *views[0])  # warmup
        t = time.perf_counter()
        loop_imgs = [bl.render_hq_scene(scene_l, p, s) for p, s in views]
        t_loop = time.perf_counter() - t
        log(f"per-view launch loop: {t_loop:.2f} s")
    finally:
        scene_l.free()
        bl.close()
    same = all(np.array_equal(x.view("<u4"), y.view("<u4")) for x, y in zip(batch_imgs, plain_imgs))
    same_loop = all(np.array_equal(x.view("<u4"), y.view("<u4")) for x, y in zip(batch_imgs, loop_imgs))
    return {"workload": "cfg2: 64 seeded views of the cfg1 world, sizes 160x90..1280x720, "
                        "spp 16..1024, 8 bounces, one batch vs 64 plain calls",
            "views": 64, "samples": samples, "steps": steps,
            "batched_msamples_per_s": samples / t_batch / 1e6,
            "plain_msamples_per_s": samples / t_plain / 1e6,
            "batched_s": t_batch, "plain_s": t_plain,
            "batched_over_plain": t_plain / t_batch, "images_identical": same,
            "launch_loop_s": t_loop, "launch_loop_msamples_per_s": samples / t_loop / 1e6,
            "batched_over_launch_loop": t_loop / t_batch, "launch_loop_images_identical": same_loop,
            "batch_passes": tb["launches"], "batch_trace_ms": tb["trace_ms"],
            "full_gpu_views": bound}


def batch_roofline(samples, blk):
    """The batch launch against HBM with the cfg1 world's algorithmic bytes per sample; no
    traffic figure (the counters were taken on the 720p frame, not on this view mix)."""
    r = roofline(1, samples, blk["batch_trace_ms"], max(1, blk["batch_passes"]), steps=0)
    if r and blk["batch_passes"] == 1:
        r["kernel"] = "k_render_tiles (one persistent launch for the batch)"
    return r


def bench_batch(args, lw, frames):
    blk = batch_block(lw, frames, steps=args.steps, spp_scale=args.spp_scale)
    samples = blk["samples"]
    return {
        "metric": METRIC,
        "value": blk["batched_msamples_per_s"],
        "unit": "Msamples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": 1,
        "ms_per_step": blk["batched_s"] * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 path arithmetic, f64 accumulation",
        "data": "synthetic (seeded value-noise voxel world, procedural sky)",
        "config": {"workload": blk["workload"], "views": 64, "samples": samples,
                   "kernel": kernel_form(), "parallelism": "1 GPU"},
        "plain_msamples_per_s": blk["plain_msamples_per_s"],
        "batched_over_plain": blk["batched_over_plain"],
        "images_identical": blk["images_identical"],
        "launch_loop_msamples_per_s": blk["launch_loop_msamples_per_s"],
        "batched_over_launch_loop": blk["batched_over_launch_loop"],
        "launch_loop_images_identical": blk["launch_loop_images_identical"],
        "batch_passes": blk["batch_passes"],
        "full_gpu_views": blk["full_gpu_views"],
        "roofline": batch_roofline(samples, blk),
        "cpu_baseline": None,
    }


def diff_block(lw, frames, steps: int, warmup: int) -> dict:
    """cfg3: per step one forward (lw_hip_render_diff) and one adjoint (lw_hip_adjoint) of the
    L2 loss against a target rendered at perturbed parameters; the adjoint covers all 115
    parameters (96 albedo factors, 16 intensities, turbidity and the sun's two angles). Kernel
    time (HIP events) and wall time of each pass."""
    from lumenwright_amd import diff
    t0 = time.perf_counter()
    prob = diff.config3_problem()
    cfg = prob.cfg
    fr = diff.frame_at(cfg, prob.scene, prob.start.sky)
    tfr = diff.frame_at(cfg, prob.scene, prob.target.sky)
    derivs = diff.sky_derivatives(cfg, prob.scene, prob.start.sky, 3e-4)
    log(f"problem built in {time.perf_counter() - t0:.1f} s: {cfg.name}")
    b = lw.HipBackend.open(0)
    sc = b.upload(fr)
    tsc = b.upload(tfr)
    try:
        target = b.render_diff(tsc, tfr.params, cfg.spp, prob.target.desc())
        desc_f = prob.start.desc()
        start_a = prob.start.with_values()
        desc_a = start_a.desc(derivs)
        for _ in range(warmup):
            img = b.render_diff(sc, fr.params, cfg.spp, desc_f)
            _, dldi = diff.l2_loss(img, target)
            b.adjoint(sc, fr.params, cfg.spp, desc_a, start_a.n_grads, dldi)
        fwd_ms = adj_ms = fwd_wall = adj_wall = 0.0
        for i in range(steps):
            t = time.perf_counter()
            img = b.render_diff(sc, fr.params, cfg.spp, desc_f)
            fwd_wall += time.perf_counter() - t
            fwd_ms += b.last_timing()["trace_ms"]
            loss, dldi = diff.l2_loss(img, target)
            t = time.perf_counter()
            g = b.adjoint(sc, fr.params, cfg.spp, desc_a, start_a.n_grads, dldi)
            adj_wall += time.perf_counter() - t
            adj_ms += b.last_timing()["trace_ms"]
            log(f"step {i + 1}/{steps}: loss {loss:.6g}, |grad| {np.linalg.norm(g):.4g}")
    finally:
        sc.free()
        tsc.free()
        b.close()
    samples = cfg.width * cfg.height * cfg.spp
    return {"workload": cfg.name + "; forward + adjoint of an L2 image loss over 115 parameters "
                                   "(96 albedo factors, 16 emitter intensities, turbidity, sun "
                                   "altitude and azimuth)",
            "width": cfg.width, "height": cfg.height, "spp": cfg.spp,
            "max_bounces": cfg.max_bounces, "steps": steps, "samples": samples,
            "forward_kernel_ms": fwd_ms / steps, "adjoint_kernel_ms": adj_ms / steps,
            "forward_wall_ms": fwd_wall / steps * 1e3, "adjoint_wall_ms": adj_wall / steps * 1e3,
            "forward_msamples_per_s": samples * steps / (fwd_ms / 1e3) / 1e6,
            "adjoint_msamples_per_s": samples * steps / (adj_ms / 1e3) / 1e6,
            "adjoint_over_forward": adj_ms / fwd_ms,
            "adjoint_over_forward_wall": adj_wall / fwd_wall}


def bench_diff(args, lw, frames):
    blk = diff_block(lw, frames, args.steps, args.warmup)
    return {
        "metric": METRIC,
        "value": blk["adjoint_msamples_per_s"],
        "unit": "Msamples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": blk["forward_wall_ms"] + blk["adjoint_wall_ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 path arithmetic, f64 accumulation and gradient sums",
        "data": "synthetic (seeded value-noise voxel world, procedural sky)",
        "config": {"workload": blk["workload"], "width": blk["width"], "height": blk["height"],
                   "spp": blk["spp"], "max_bounces": blk["max_bounces"], "parallelism": "1 GPU"},
        "value_is": "adjoint pass Msamples/s (kernel time)",
        "forward_msamples_per_s": blk["forward_msamples_per_s"],
        "adjoint_over_forward": blk["adjoint_over_forward"],
        "adjoint_over_forward_wall": blk["adjoint_over_forward_wall"],
        "forward_kernel_ms": blk["forward_kernel_ms"],
        "adjoint_kernel_ms": blk["adjoint_kernel_ms"],
        "roofline": None,
        "cpu_baseline": None,
    }


def fit_block(lw, frames, steps: int, warmup: int, fit_steps: int = 0, spp: int = 128) -> dict:
    """cfg4: the batched inverse fit. 32 seeded views of the cfg1 world (640x360, `spp`, 6
    bounces); targets rendered with seeded ground-truth factors (192 albedo factors of the 64
    block types, 128 emitter-intensity groups); Adam on the device from factors of 1. A step is
    one batched forward, the loss and dL/dI, one batched adjoint (path replay) and the update,
    all on the device. `warmup` untimed steps, then `steps` timed ones (of max(steps,
    fit_steps) in all)."""
    from lumenwright_amd import diff
    t0 = time.perf_counter()
    bp = diff.config4_problem(spp=spp)
    log(f"problem built in {time.perf_counter() - t0:.1f} s: {len(bp.views)} views, "
        f"{bp.start.n_palette} palettes, {bp.start.n_intensity} intensity groups")
    b = lw.HipBackend.open(0)
    sc = b.upload(bp.base)
    fit = None
    try:
        t = time.perf_counter()
        targets = b.render_diff_batch(sc, bp.views, bp.truth.desc())
        log(f"targets rendered in {time.perf_counter() - t:.1f} s")
        fit = b.fit(sc, bp.views, bp.start.desc(), targets, lr=1e-2)
        e_start = diff.parameter_error(bp.start, bp.truth)
        for i in range(warmup):
            st = fit.step()
            log(f"warmup {i + 1}: {st}")
        done = []
        t = time.perf_counter()
        n_steps = max(steps, fit_steps)
        for i in range(n_steps):
            st = fit.step()
            done.append(st)
            if i < steps or (i + 1) % 10 == 0:
                log(f"step {i + 1}/{n_steps}: loss {st['loss']:.6g}, forward {st['forward_ms']:.0f} ms, "
                    f"adjoint {st['adjoint_ms']:.0f} ms, wall {st['wall_ms']:.0f} ms")
        wall = time.perf_counter() - t
        pal, inten, _ = fit.params()
        e_end = diff.parameter_error(bp.start.with_values(pal, inten), bp.truth)
    finally:
        if fit is not None:
            fit.free()
        sc.free()
        b.close()
    timed = done[:steps]
    fwd = sum(s["forward_ms"] for s in timed)
    adj = sum(s["adjoint_ms"] for s in timed)
    step_wall = sum(s["wall_ms"] for s in timed) / len(timed)
    samples = timed[0]["samples"]
    return {"workload": "cfg4: batched inverse fit, 32 seeded views of the cfg1 world, "
                        f"640x360, {spp} spp, 6 bounces; 192 albedo factors and 128 "
                        "emitter-intensity groups; Adam lr 1e-2 on the device",
            "views": len(bp.views), "spp": spp, "samples": samples, "steps": steps,
            "step_wall_ms": step_wall,
            "forward_kernel_ms": fwd / len(timed), "adjoint_kernel_ms": adj / len(timed),
            "forward_msamples_per_s": samples * len(timed) / (fwd / 1e3) / 1e6,
            "adjoint_msamples_per_s": samples * len(timed) / (adj / 1e3) / 1e6,
            "adjoint_over_forward": adj / fwd,
            "fit_steps": n_steps, "loss_first": done[0]["loss"], "loss_last": done[-1]["loss"],
            "param_error_start": {"albedo": e_start[0], "intensity": e_start[1]},
            "param_error_end": {"albedo": e_end[0], "intensity": e_end[1]},
            "fit_wall_s": wall}


def bench_fit(args, lw, frames):
    blk = fit_block(lw, frames, args.steps, args.warmup, args.fit_steps, args.spp or 128)
    return {
        "metric": METRIC,
        "value": blk["samples"] / (blk["step_wall_ms"] / 1e3) / 1e6,
        "unit": "Msamples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": blk["step_wall_ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 path arithmetic, f64 accumulation and gradient sums",
        "data": "synthetic (seeded value-noise voxel world, game dusk sky maps)",
        "config": {"workload": blk["workload"], "views": blk["views"], "spp": blk["spp"],
                   "parallelism": "1 GPU"},
        "value_is": "view samples per second of a whole fit step (forward, loss, adjoint, update)",
        **{k: blk[k] for k in ("forward_msamples_per_s", "adjoint_msamples_per_s",
                               "adjoint_over_forward", "fit_steps", "loss_first", "loss_last",
                               "param_error_start", "param_error_end", "fit_wall_s")},
        "roofline": None,
        "cpu_baseline": None,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", type=int, default=1, choices=[0, 1, 2, 3, 4])
    ap.add_argument("--fit-steps", type=int, default=0,
                    help="config 4: run this many fit steps in all (the first --steps are timed)")
    ap.add_argument("--spp", type=int, default=None, help="override the workload's spp")
    ap.add_argument("--spp-scale", type=float, default=1.0, help="config 2: scale every view's spp")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-metric-parts", action="store_true",
                    help="config 1: leave out the batched (cfg2) and adjoint (cfg3, cfg4) blocks")
    ap.add_argument("--copy-probe", type=int, default=0, help="also time a copy of this many bytes")
    ap.add_argument("--autotune", action="store_true",
                    help="pick the launch shape by timing probes first (lw_hip_autotune)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: --gpus {args.gpus} is not supported (the path "
              "shards by sample range or tile with one f64 reduction, which this build does "
              "not run)", file=sys.stderr)
        sys.exit(2)

    import lumenwright_amd as lw
    from lumenwright_amd import frames

    if args.config == 2:
        line = bench_batch(args, lw, frames)
    elif args.config == 3:
        line = bench_diff(args, lw, frames)
    elif args.config == 4:
        line = bench_fit(args, lw, frames)
    else:
        line = bench_plain(args, lw, frames)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
