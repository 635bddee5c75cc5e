"""Benchmark: 1080p frames/s at 1,048,576 synthetic Gaussians (SH degree 3).

One step = one full frame of the render path on one MI355X: project, depth
argsort, gather scan, isect emission, tile argsort, tile ranges, RGB
composite (f32 RGBA + RGBA8).  Inputs are resident in HBM before the timed
region.  Prints one JSON line (see the repository's benchmark contract).
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=5)
    p.add_argument("--n", type=int, default=1 << 20)
    p.add_argument("--width", type=int, default=1920)
    p.add_argument("--height", type=int, default=1080)
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    p.add_argument("--cpu-threads", type=int, default=0)
    return p.parse_args()


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: every box in this setup has a single MI355X; "
              "multi-GPU runs are not part of this build", file=sys.stderr)
        sys.exit(2)
    from splatfield_amd import Context
    from splatfield_amd import render as R
    from splatfield_amd import scene as S

    n, w, h = args.n, args.width, args.height
    attrs, sh, k = S.synthetic_scene(n, 3, seed=0)
    fy = np.float32(math.pi / 3)
    fx = np.float32(S.fov_for_aspect(w, h, fy))
    cam = R.Camera.look_at((0.0, 0.0, 3.0), (0.0, 0.0, 0.0), fov=(fx, fy))

    ctx = Context(0)
    splats = R.Splats(ctx, attrs, sh)
    scratch = R.RenderScratch(ctx, n, w, h)
    view = cam.view(w, h)

    def step():
        isects = R.prepare_isects(splats, scratch, view=view, w=w, h=h)
        R.finalize_rgb(scratch, isects)
        return isects

    isects = step()
    ctx.sync()
    raw, vis, ni = isects.counts()
    scratch.reserve_isects(raw)
    for _ in range(args.warmup):
        step()
    ctx.sync()
    raw, vis, ni = isects.counts()
    assert ni == raw, "isect capacity must be steady before timing"

    ctx.timer_start()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    gpu_ms = ctx.timer_stop_ms()
    wall = time.perf_counter() - t0
    ms_per_step = gpu_ms / args.steps

    # Algorithmic bytes of the forward frame (two-stage sort at 8-bit digits).
    tile_bits = int(math.ceil(math.log2(((w + 15) // 16) * ((h + 15) // 16))))
    b_fwd = (n * (44 + 4 * 3 * k) + 56 * vis + 16 * vis + 8 * ni + 4 * 16 * vis +
             math.ceil(tile_bits / 8) * 16 * ni + 40 * ni + 16 * w * h)
    achieved = b_fwd / (ms_per_step * 1e-3) / 1e9

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(attrs, sh, n, k, view, w, h, args.cpu_threads)

    line = {
        "metric": "frames/s at 1M Gaussians 1080p (forward render)",
        "value": round(1000.0 / ms_per_step, 2),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded splats: means U[-1,1]^3, log-scales N(-4,0.5), SH3)",
        "config": {"workload": "1080p forward frame, 1,048,576 Gaussians SH3",
                   "gaussians": n, "width": w, "height": h, "isects": ni, "visible": vis},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": None, "kernel": "whole frame"},
        "cpu_baseline": cpu,
        "wall_s": round(wall, 3),
    }
    print(json.dumps(line))
    ctx.close()


def cpu_baseline(attrs, sh, n, k, view, w, h, threads):
    """The oracle's C restatement of the same frame, timed on the host."""
    from oracle import oracle as O
    cores = threads or min(16, os.cpu_count() or 1)
    ov = O.View.from_buffer_copy(bytes(view))
    t0 = time.perf_counter()
    O.render(ov, attrs, sh, n, k, w, h, nthreads=cores)
    dt = time.perf_counter() - t0
    return {"value": round(1.0 / dt, 4), "unit": "frames/s", "cores": cores, "kind": "port",
            "sample": "one full 1080p forward frame of the same scene (project, sorts, map, "
                      f"composite) by the oracle, {dt:.1f} s"}


if __name__ == "__main__":
    main()
