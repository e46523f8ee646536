#!/usr/bin/env python3
"""Benchmark: Msamples/s of the HQ path tracer on one MI355X.

A step is one render_hq of the workload (default: config 1, 1280 x 720 at 256 spp with 8
bounces over the 24 x 24-chunk world with water and emitters, game dusk sky) with the
scene resident in HBM. W untimed warmup steps, then K timed steps; prints one JSON line.

Besides the throughput the line carries:
  roofline      the path-tracing kernel's algorithmic scene bytes per launch (per-sample
                bytes counted once by the instrumented oracle, profiles/b_alg.json, times
                the samples of a launch) over its mean launch time (HIP events on the
                library's stream), against the 8 TB/s HBM peak;
  cpu_baseline  the oracle (the C++ restatement of the reference's CPU backend) on the
                host's cores over a bounded slice of the same frame (the first S samples of
                every pixel), timed in this run.
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


def host_cores() -> int:
    n = len(os.sched_getaffinity(0))
    cap = os.environ.get("OMP_NUM_THREADS")
    if cap and cap.isdigit():
        n = min(n, int(cap))
    return max(1, n)


def cpu_baseline(frame, seconds: float):
    """The oracle's CpuBackend::render_hq over the frame's first spp samples, doubling spp
    until one render takes at least `seconds / 4`, then repeating to fill `seconds`."""
    from oracle import ffi
    ffi.build()
    from lumenwright_amd.abi import HqFrame
    cores = host_cores()
    v = HqFrame(frame.scene, frame.params, 1, frame.sky).view()
    w, h = int(frame.params["frame"][0]), int(frame.params["frame"][1])
    spp, total_s, total_n = 1, 0.0, 0
    while True:
        t = time.perf_counter()
        ffi.render_hq(v, spp, threads=cores)
        dt = time.perf_counter() - t
        total_s += dt
        total_n += w * h * spp
        if total_s >= seconds or spp >= frame.spp:
            break
        if dt < seconds / 4:
            spp = min(spp * 2, frame.spp)
    return {"value": total_n / total_s / 1e6, "unit": "Msamples/s", "cores": cores,
            "kind": "port",
            "sample": f"{w}x{h}, the first {spp} samples of every pixel (and smaller passes "
                      f"before it), {total_n} samples in {total_s:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", type=int, default=1)
    ap.add_argument("--spp", type=int, default=None, help="override the workload's spp")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: --gpus {args.gpus} is not supported (the path "
              "shards by sample range or tile with one f64 reduction, which this build does "
              "not run)", file=sys.stderr)
        sys.exit(2)

    import lumenwright_amd as lw
    from lumenwright_amd import frames

    cfg = frames.CONFIGS[args.config]
    if args.spp:
        cfg = frames.config(args.config, spp=args.spp)
    t0 = time.perf_counter()
    frame = frames.make_frame(cfg)
    print(f"# frame built in {time.perf_counter() - t0:.1f} s: {cfg.name}", file=sys.stderr,
          flush=True)

    b = lw.HipBackend.open(0)
    scene = b.upload(frame)
    w, h, spp = cfg.width, cfg.height, cfg.spp
    for i in range(args.warmup):
        b.render_hq_scene(scene, frame.params, spp)
        print(f"# warmup {i + 1}/{args.warmup}: {b.last_timing()}", file=sys.stderr, flush=True)
    trace_ms, launches, acc_ms = 0.0, 0, 0.0
    t = time.perf_counter()
    for i in range(args.steps):
        img = b.render_hq_scene(scene, frame.params, spp)
        tm = b.last_timing()
        trace_ms += tm["trace_ms"]
        acc_ms += tm["accumulate_ms"]
        launches += tm["launches"]
        print(f"# step {i + 1}/{args.steps}: {tm}", file=sys.stderr, flush=True)
    elapsed = time.perf_counter() - t
    samples = w * h * spp * args.steps
    value = samples / elapsed / 1e6
    assert np.isfinite(img).all()

    roofline = None
    b_alg_path = ROOT / "profiles" / "b_alg.json"
    if b_alg_path.exists():
        b_alg = json.loads(b_alg_path.read_text()).get(f"config{args.config}")
        if b_alg:
            per_launch_samples = samples / launches
            mean_launch_s = trace_ms / 1e3 / launches
            bytes_per_launch = b_alg["bytes_per_sample"] * per_launch_samples
            achieved = bytes_per_launch / mean_launch_s / 1e9
            roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                        "kernel": "k_render", "mean_launch_ms": trace_ms / launches,
                        "bytes_per_sample": b_alg["bytes_per_sample"]}

    cpu = None if args.no_cpu_baseline else cpu_baseline(frame, args.cpu_seconds)
    line = {
        "metric": "Msamples/s at 720p 8-bounce hq",
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
                   "max_bounces": cfg.max_bounces, "parallelism": "1 GPU"},
        "kernel_ms_per_step": trace_ms / args.steps,
        "accumulate_ms_per_step": acc_ms / args.steps,
        "launches_per_step": launches / args.steps,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    if cpu:
        line["speedup_vs_cpu_baseline"] = value / cpu["value"]
    print(json.dumps(line), flush=True)
    scene.free()
    b.close()


if __name__ == "__main__":
    main()
