"""Benchmark: 1080p forward+backward frames/s at 1,048,576 Gaussians (SH3).

One step = one differentiable frame on one MI355X: project, depth argsort,
gather scan, isect emission, tile argsort, tile ranges, RGB composite
(f32 RGBA + per-pixel final T / last contributor), L2 loss gradient against a
target image, raster adjoint and projection adjoint (gradients w.r.t. means,
quaternions, log-scales, opacity logits and SH).  Inputs and the target are
resident in HBM before the timed region.  Prints one JSON line.
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
    p.add_argument("--cpu-seconds", type=float, default=10.0)
    p.add_argument("--no-extra", action="store_true",
                   help="skip the EIG-round and standalone-sort measurements")
    p.add_argument("--eig-views", type=int, default=20)
    p.add_argument("--eig-res", type=int, default=512)
    p.add_argument("--eig-reps", type=int, default=5)
    p.add_argument("--leg", choices=["eig", "sort", "c4"], default=None,
                   help="run only this extra measurement (no frame), print its JSON")
    p.add_argument("--calibrate-copy", action="store_true",
                   help="only run a 256 MiB device copy (known traffic, for counter scaling)")
    return p.parse_args()


def calibrate_copy(ctx):
    from splatfield_amd.runtime import GpuTensor
    from splatfield_amd._lib import lib
    n = 64 << 20  # u32 -> 256 MiB
    src = GpuTensor.empty(ctx, n)
    dst = GpuTensor.empty(ctx, n)
    src.fill_u32(7)
    ctx.sync()
    ctx.timer_start()
    for _ in range(5):
        ctx.check(lib().sfk_copy_u32(ctx.ptr, dst.ptr, src.ptr, n), "copy")
    ms = ctx.timer_stop_ms() / 5
    print(json.dumps({"copy_bytes_read": n * 4, "copy_bytes_written": n * 4,
                      "ms": round(ms, 4), "gbs": round(2 * n * 4 / (ms * 1e-3) / 1e9, 1)}))


def stage(msg):
    """Progress line on stderr (the JSON result is the last stdout line)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def eig_round_bytes(n, views, batched):
    """Algorithmic bytes of an EIG round over 512^2 views at 16-px tiles
    (10 tile bits), per view: attributes 44N (once when batched), projected
    rows 48V, depth sort 4 x 16V, emission read 16V, isects 8I, tile sort
    16 x 2 x I, raster 28I, responsibility clear + write 8N, EIG read 12N.
    `views` holds (isects, visible) per view."""
    total = 44 * n if batched else 0
    for ni, vis in views:
        total += (0 if batched else 44 * n) + 48 * vis + 64 * vis + 16 * vis + 8 * ni \
            + 16 * 2 * ni + 28 * ni + 8 * n + 12 * n
    return total


def bench_eig_round(ctx, args):
    """The B3-Seg evidence round on the 1M-splat SH3 scene: a Beta(1,1)
    prior updated by one observation (the footprint of the splats within 0.3
    of a seeded centre, seen from (0, 0, 3) at 512^2), then 20 Fibonacci
    candidate views at 512^2 scored by analytic f64 EIG.  Plain = the
    reference's select_view loop (per view a responsibility render and a
    synchronous EIG) at each tile width; batched = one hipGraph replay at the
    autotuned tile width x digit width.  Ratios compare equal tile widths
    (the evidence is then bit-identical)."""
    from splatfield_amd import render as R
    from splatfield_amd import scene as S
    from splatfield_amd import seg as G
    from splatfield_amd.runtime import GpuTensor
    n, res, nv = args.n, args.eig_res, args.eig_views
    attrs, sh, k = S.synthetic_scene(n, 3, seed=0)
    splats = R.Splats(ctx, attrs, sh)
    cam0 = R.Camera.look_at((0.0, 0.0, 3.0), (0.0, 0.0, 0.0))
    mask, subset = G.footprint_mask(ctx, attrs, sh, n, k, cam0, res, res, seed=0)
    state = G.observe_bootstrap(splats, cam0, res, mask)
    cams = G.sphere_cameras(nv, radius=3.0)
    sizes = [(res, res)] * nv
    resp = GpuTensor.empty(ctx, n)
    plain = {}
    for tw in (8, 16, 32):
        stage(f"eig round: plain loop, tile width {tw}")
        scratch = R.RenderScratch(ctx, n, res, res, tile_w=tw)
        scratch.reserve_isects(1 << 24)
        G.eig_round_plain(splats, scratch, cams, sizes, state, resp)  # warm
        t = []
        for _ in range(args.eig_reps):
            t0 = time.perf_counter()
            eig_p, best_p = G.eig_round_plain(splats, scratch, cams, sizes, state, resp)
            t.append(time.perf_counter() - t0)
        plain[tw] = (1e3 * float(np.median(t)), eig_p, best_p, resp.read())
        del scratch

    def timed_round(tw, db):
        rnd = G.EigRound(splats, cams, sizes, tile_w=tw, digit_bits=db)
        rnd.run(state, use_graph=True)
        tb = []
        for _ in range(args.eig_reps):
            t0 = time.perf_counter()
            eig_b, best_b = rnd.run(state, use_graph=True)
            tb.append(time.perf_counter() - t0)
        return rnd, 1e3 * float(np.median(tb)), eig_b, best_b

    stage("eig round: autotune")
    (tw_b, db_b), table = G.autotune_round(splats, cams, sizes, state, reps=3)
    rnd, batch_ms, eig_b, best_b = timed_round(tw_b, db_b)
    p_ms, eig_p, best_p, last_resp = plain[tw_b]
    exact = bool(np.array_equal(rnd.resp(nv - 1)[0], last_resp))
    cnt = rnd.counters()
    per_view = [(int(cnt[2 * v]), int(cnt[2 * v + 1])) for v in range(nv)]
    del rnd
    out = {"workload": "Beta(1,1) prior + one footprint-mask update, 20 Fibonacci views "
                       "(radius 3, fov 0.8) at 512^2, 1M SH3 splats",
           "views": nv, "res": res, "mask_fg_pixels": int(mask.sum()),
           "target_splats": int(subset.sum()),
           "plain_ms_by_tile_w": {str(t_): round(v_[0], 3) for t_, v_ in plain.items()},
           "batched_ms": round(batch_ms, 3), "batched_tile_w": tw_b, "batched_digit_bits": db_b,
           "autotune_ms": {f"tile{t_}_digit{d_}": round(v_, 3)
                           for (t_, d_), v_ in table.items()},
           "ratio": round(batch_ms / p_ms, 4),
           "ratio_vs_fastest_plain": round(batch_ms / min(v_[0] for v_ in plain.values()), 4),
           "views_per_s": round(nv / (batch_ms * 1e-3), 1),
           "best_plain": best_p, "best_batched": best_b,
           "eig_max_abs_diff": float(np.max(np.abs(eig_b - eig_p))),
           "last_view_evidence_bit_exact": exact,
           "isects_per_view": [v_[0] for v_ in per_view],
           "visible_per_view": [v_[1] for v_ in per_view]}
    # The reference's 16-px tiles on both sides, and the round's HBM
    # fraction on the algorithmic bytes of that configuration.
    if tw_b == 16:
        ms16, db16, views16 = batch_ms, db_b, per_view
    else:
        r16, ms16, _, _ = timed_round(16, 8)
        c16 = r16.counters()
        views16 = [(int(c16[2 * v]), int(c16[2 * v + 1])) for v in range(nv)]
        db16 = 8
        del r16
    b16 = eig_round_bytes(n, views16, True)
    p16 = eig_round_bytes(n, views16, False)
    out["tile16"] = {"batched_ms": round(ms16, 3), "digit_bits": db16,
                     "plain_ms": round(plain[16][0], 3), "ratio": round(ms16 / plain[16][0], 4),
                     "batched_algorithmic_bytes": b16, "plain_algorithmic_bytes": p16,
                     "batched_gbs": round(b16 / (ms16 * 1e-3) / 1e9, 1),
                     "batched_frac": round(b16 / (ms16 * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                     "plain_gbs": round(p16 / (plain[16][0] * 1e-3) / 1e9, 1)}
    return out


def bench_c4(ctx, args):
    """Mixed-resolution batched round: 2,097,152 splats (SH3), 64
    Fibonacci-sphere cameras cycling 256^2 / 512^2 / 1024^2.  The autotuner
    picks tile width x digit width, the chosen round is timed as graph
    replays, and the plain loop (one scratch per resolution, isect capacity
    reserved) is timed beside it."""
    from splatfield_amd import render as R
    from splatfield_amd import scene as S
    from splatfield_amd import seg as G
    n = 1 << 21
    attrs, sh, k = S.synthetic_scene(n, 3, seed=0)
    splats = R.Splats(ctx, attrs, sh)
    cams = G.sphere_cameras(64, radius=3.0)
    res = (256, 512, 1024)
    sizes = [(res[i % 3], res[i % 3]) for i in range(64)]
    rng = np.random.default_rng(2)
    state = G.BetaState(ctx, (1 + rng.random(n) * 4).astype(np.float32),
                        (1 + rng.random(n) * 4).astype(np.float32))
    stage("c4 round: autotune")
    best, table = G.autotune_round(splats, cams, sizes, state, reps=2)
    rnd = G.EigRound(splats, cams, sizes, tile_w=best[0], digit_bits=best[1])
    rnd.run(state, use_graph=True)
    tb = []
    for _ in range(3):
        t0 = time.perf_counter()
        eig_b, best_b = rnd.run(state, use_graph=True)
        tb.append(time.perf_counter() - t0)
    isects = int(rnd.counters()[-1])
    del rnd
    scratches = {}
    acc = G.Accumulators(ctx, n)

    def plain():
        e = []
        for cam, (w, h) in zip(cams, sizes):
            sc = scratches.get(w)
            if sc is None:
                sc = scratches[w] = R.RenderScratch(ctx, n, w, h)
                sc.reserve_isects(1 << 25)
            G.render_responsibility(splats, sc, acc, cam, w, h)
            e.append(state.eig(acc))
        return e

    stage("c4 round: plain loop")
    plain()
    tp = []
    for _ in range(2):
        t0 = time.perf_counter()
        eig_p = plain()
        tp.append(time.perf_counter() - t0)
    batch_ms, plain_ms = 1e3 * float(np.median(tb)), 1e3 * float(np.median(tp))
    return {"gaussians": n, "views": 64, "res": list(res),
            "autotune_ms": {f"tile{t}_digit{d}": round(v, 3) for (t, d), v in table.items()},
            "tile_w": best[0], "digit_bits": best[1], "batched_ms": round(batch_ms, 3),
            "plain_ms": round(plain_ms, 3), "ratio": round(batch_ms / plain_ms, 4),
            "views_per_s": round(64 / (batch_ms * 1e-3), 1), "isects_total": isects,
            "same_best_view": bool(int(np.argmax(eig_p)) == best_b),
            "eig_max_rel_diff": float(np.max(np.abs(eig_b - eig_p) / np.abs(eig_p)))}


def bench_sort(ctx, args):
    """Standalone stable argsort of 16,777,216 uniform u64 keys + u32 index
    values (HIP events around the sort launches only)."""
    from splatfield_amd import sort as SO
    from splatfield_amd.runtime import GpuTensor
    n = 1 << 24
    keys = np.random.default_rng(0).integers(0, 2**64 - 1, n, dtype=np.uint64)
    kt = GpuTensor.from_numpy(ctx, keys)
    vt = GpuTensor.from_numpy(ctx, np.arange(n, dtype=np.uint32))
    s = SO.RadixScratch(ctx, n, 8, 8)
    s.argsort(kt, vt, n, 64)
    ctx.sync()
    reps = 5
    ctx.timer_start()
    for _ in range(reps):
        s.argsort(kt, vt, n, 64)
    ms = ctx.timer_stop_ms() / reps
    # The fixed bytes of the reference's eight 8-bit passes, whatever the
    # strategy; and the bucket finish's own minimal traffic: two passes
    # (8 B key histogram read, 12 B read + 12 B written), the bucket bounds
    # (8 B), the on-chip bucket sort (12 B read, 12 B written).
    algo = n * 8 + 8 * 2 * (8 + 4) * n
    implemented = n * (2 * (8 + 24) + 8 + 24)
    return {"keys": n, "ms": round(ms, 4), "gkeys_per_s": round(n / (ms * 1e-3) / 1e9, 3),
            "achieved_gbs": round(algo / (ms * 1e-3) / 1e9, 1),
            "frac": round(algo / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "strategy": "two 8-bit passes on the top 16 bits, then every 16-bit group sorted on "
                        "chip (one wavefront each); the call waits once for a group-size check",
            "implemented_bytes": implemented,
            "implemented_frac": round(implemented / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}


def bench_paint(ctx, splats, scratch, cam, w, h, n, reps=20):
    """Posterior-mean render (src/seg/paint.rs:22-47) against the colour
    render on the frame's scene.  The paint rides in the projection, which
    then reads 8 bytes of (a, b) per splat instead of 4K of SH."""
    import ctypes as C
    from splatfield_amd._lib import lib
    from splatfield_amd.runtime import GpuTensor
    rng = np.random.default_rng(2)
    a = GpuTensor.from_numpy(ctx, (1 + rng.gamma(1.0, 2.0, n)).astype(np.float32))
    b = GpuTensor.from_numpy(ctx, (1 + rng.gamma(1.0, 2.0, n)).astype(np.float32))
    c = cam.to_c()

    def colours():  # no bitmap readback inside the timed region
        ctx.check(lib().sfk_render_with(ctx.ptr, splats.ptr, scratch.ptr, C.byref(c), w, h,
                                        None, None), "render_with")

    def posterior():
        ctx.check(lib().sfk_render_posterior(ctx.ptr, splats.ptr, scratch.ptr, a.ptr, b.ptr, n,
                                             C.byref(c), w, h, None, None), "render_posterior")
    out = {}
    for name, fn in (("render_with_ms", colours), ("render_posterior_ms", posterior)):
        fn()
        ctx.timer_start()
        for _ in range(reps):
            fn()
        out[name] = round(ctx.timer_stop_ms() / reps, 4)
    for name, fn in (("project_colour_ms", colours), ("project_posterior_ms", posterior)):
        ctx.prof_enable(True)
        for _ in range(reps):
            fn()
        cnt, ms = ctx.prof_report()["project"]
        ctx.prof_enable(False)
        out[name] = round(ms / cnt, 5)
    return out


def bench_ply(ctx, attrs, sh, n, reps=10):
    """PLY ingestion (src/ply.rs:9-112) of the frame's scene: the device
    transpose of the file's vertex rows (inputs resident), its algorithmic
    bytes the rows read plus the planes written."""
    import ctypes as C
    from splatfield_amd import render as R
    from splatfield_amd import scene as S
    from splatfield_amd._lib import lib
    from splatfield_amd.runtime import GpuTensor
    data = S.write_ply(attrs, sh, n)
    lay = R.parse_ply_header(data)
    rows = np.frombuffer(data, np.uint8, lay.n * lay.stride * 4, lay.data_offset)
    d_rows = GpuTensor.from_numpy(ctx, rows)
    d_attrs = GpuTensor.empty(ctx, 11 * n, np.float32)
    d_sh = GpuTensor.empty(ctx, lay.sh_planes * n, np.float32)

    def dec():
        ctx.check(lib().sfk_ply_decode(ctx.ptr, C.byref(lay), d_rows.ptr, d_attrs.ptr,
                                       d_sh.ptr), "ply_decode")
    dec()
    ctx.timer_start()
    for _ in range(reps):
        dec()
    ms = ctx.timer_stop_ms() / reps
    algo = n * lay.stride * 4 + n * (11 + lay.sh_planes) * 4
    return {"splats": n, "row_floats": int(lay.stride), "ms": round(ms, 4),
            "achieved_gbs": round(algo / (ms * 1e-3) / 1e9, 1),
            "frac": round(algo / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}


def bench_sog(ctx, n, reps=10):
    """SOG decode (src/sog.rs:108-195) of n SH3 splats from device-resident
    RGBA8 sheets (synthetic texels, 1024-texel rows, a 4096-palette centroid
    sheet): 24 bytes of texels read and 236 bytes of planes written per
    splat (palette gathers hit cache)."""
    import ctypes as C
    from splatfield_amd import _lib
    from splatfield_amd._lib import lib
    from splatfield_amd.runtime import GpuTensor
    rng = np.random.default_rng(3)
    sheets = {k: rng.integers(0, 256, 4 * n, dtype=np.uint8)
              for k in ("lo", "hi", "scales", "quats", "sh0", "labels")}
    sheets["quats"][3::4] = rng.integers(252, 256, n)
    sheets["labels"][1::4] = rng.integers(0, 16, n)  # labels < 4096
    cw, cpc = 1020, 15
    sheets["centroids"] = rng.integers(0, 256, 4 * cw * 61, dtype=np.uint8)
    dev = {k: GpuTensor.from_numpy(ctx, v) for k, v in sheets.items()}
    cb = GpuTensor.from_numpy(ctx, rng.normal(0, 1, 768).astype(np.float32))
    m = _lib.SogMeta()
    m.n, m.sh_per_ch, m.centroid_w = n, 16, cw
    m.centroid_px = cw * 61
    m.palette_count = m.centroid_px // cpc
    m.mins[:] = [-2.0, -2.0, -2.0]
    m.maxs[:] = [2.0, 2.0, 2.0]
    attrs = GpuTensor.empty(ctx, 11 * n, np.float32)
    sh = GpuTensor.empty(ctx, 48 * n, np.float32)
    status = GpuTensor.empty(ctx, 2, np.uint32)

    def dec():
        ctx.check(lib().sfk_sog_decode(ctx.ptr, C.byref(m), dev["lo"].ptr, dev["hi"].ptr,
                                       dev["scales"].ptr, dev["quats"].ptr, dev["sh0"].ptr,
                                       dev["labels"].ptr, dev["centroids"].ptr, cb.ptr,
                                       attrs.ptr, sh.ptr, status.ptr), "sog_decode")
    dec()
    ctx.timer_start()
    for _ in range(reps):
        dec()
    ms = ctx.timer_stop_ms() / reps
    assert list(status.read()) == [0xFFFFFFFF, 0xFFFFFFFF]
    algo = n * (24 + 44 + 192)
    return {"splats": n, "ms": round(ms, 4),
            "achieved_gbs": round(algo / (ms * 1e-3) / 1e9, 1),
            "frac": round(algo / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}


def kernel_bytes(n, k, vis, ni, px, pairs):
    """Algorithmic HBM bytes per launch of each step of the frame: the reads
    and writes the implemented pipeline must make (not counting cache
    reuse).  Radix passes move key + value and read the key once more for
    the histogram: 28 B per element with 8-byte values."""
    return {
        "project": n * (44 + 4 * 3 * k) + 56 * vis,
        "radix_pass_u32": None,  # depth and pair passes differ; see main()
        # (splat, bbox) per rank read; depth order and offsets written;
        # (supertile, splat, bbox) pairs written
        "bin_scan": 16 * n + 12 * pairs,
        "bin_count": 8 * pairs + 256 * (pairs // 512),
        # pairs read again; splat id written per isect (the frame's scratch
        # writes no tile ids)
        "bin_place": 8 * pairs + 4 * ni,
        "raster_rgb": 40 * ni + 16 * px,
        "loss_l2": 48 * px,
        "raster_bwd": 40 * ni + 32 * px + 36 * vis,
        "project_bwd": 2 * n * (44 + 4 * 3 * k) + 36 * vis,
    }


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: every box in this setup has a single MI355X; "
              "multi-GPU runs are not part of this build", file=sys.stderr)
        sys.exit(2)
    from splatfield_amd import Context
    from splatfield_amd import render as R
    from splatfield_amd import scene as S
    from splatfield_amd.runtime import GpuTensor

    if args.calibrate_copy:
        ctx = Context(0)
        calibrate_copy(ctx)
        ctx.close()
        return
    if args.leg:
        ctx = Context(0)
        fn = {"eig": bench_eig_round, "sort": bench_sort, "c4": bench_c4}[args.leg]
        print(json.dumps({args.leg: fn(ctx, args)}))
        ctx.close()
        return
    n, w, h = args.n, args.width, args.height
    px = w * h
    stage(f"frame: {n} splats, {w}x{h}, forward+backward")
    attrs, sh, k = S.synthetic_scene(n, 3, seed=0)
    rng = np.random.default_rng(1)
    pa = (attrs + rng.normal(0, 0.01, attrs.size)).astype(np.float32)
    ps = (sh + rng.normal(0, 0.01, sh.size)).astype(np.float32)
    fy = np.float32(math.pi / 3)
    fx = np.float32(S.fov_for_aspect(w, h, fy))
    cam = R.Camera.look_at((0.0, 0.0, 3.0), (0.0, 0.0, 0.0), fov=(fx, fy))

    ctx = Context(0)
    view = cam.view(w, h)
    # Target image: the same scene with parameters perturbed by N(0, 0.01).
    sp_t = R.Splats(ctx, pa, ps)
    sc_t = R.RenderScratch(ctx, n, w, h)
    sc_t.reserve_isects(16 << 20)
    it = R.prepare_isects(sp_t, sc_t, view=view, w=w, h=h)
    R.finalize_rgb(sc_t, it)
    target = GpuTensor.from_numpy(ctx, sc_t.read("rgba"))
    del it, sc_t, sp_t

    splats = R.Splats(ctx, attrs, sh)
    scratch = R.RenderScratch(ctx, n, w, h)
    # The frame never reads the isect arrays: no per-isect tile ids, and the
    # rasterizers take each tile's isects from the sorted supertile pairs,
    # placing only the ones they consume (the backward reads those).
    scratch.set_tile_ids(False)
    scratch.set_pair_raster(True)
    loss = GpuTensor.empty(ctx, 1, np.float64)
    g_attrs = GpuTensor.empty(ctx, n * 11, np.float32)
    g_sh = GpuTensor.empty(ctx, n * 3 * k, np.float32)
    rgba_ptr = None

    def step():
        isects = R.prepare_isects(splats, scratch, view=view, w=w, h=h)
        R.finalize_rgb(scratch, isects)
        R.render_backward_l2(splats, scratch, isects, rgba_ptr, target.ptr, g_attrs, g_sh, loss)
        return isects

    rgba_ptr = scratch.device("rgba")
    isects = step()
    ctx.sync()
    raw, vis, ni = isects.counts()
    scratch.reserve_isects(raw)
    rgba_ptr = scratch.device("rgba")
    for _ in range(max(args.warmup, 1)):
        isects = step()
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
    loss_v = float(loss.read()[0])

    # Same steps again with per-launch events: the kernel breakdown.
    ctx.prof_enable(True)
    for _ in range(args.steps):
        step()
    prof = ctx.prof_report()
    ctx.prof_enable(False)

    pairs = int(scratch.read("counters")[1])
    kb = kernel_bytes(n, k, vis, ni, px, pairs)
    # Radix passes per frame: the depth sort (three passes on rebased 24-bit
    # keys where the view's depth window allows, else four) over all n keys,
    # then one pass over the supertile pairs.
    radix_launches = prof.get("radix_pass_u32", (0, 0.0))[0] / args.steps
    depth_passes = max(int(round(radix_launches)) - 1, 0)
    kb["radix_pass_u32"] = (depth_passes * 28 * n + 28 * pairs) / max(depth_passes + 1, 1)
    per_kernel = {}
    for name, (cnt, ms) in prof.items():
        per_launch = ms / max(cnt, 1)
        entry = {"launches_per_step": cnt / args.steps, "ms_per_launch": round(per_launch, 5)}
        if kb.get(name):
            entry["gbs"] = round(kb[name] / (per_launch * 1e-3) / 1e9, 1)
        per_kernel[name] = entry
    # Dominant kernel: the single kernel with the largest time per step
    # (radix_pass_* scopes span three kernels and two workloads).
    single = {k_: v_ for k_, v_ in prof.items() if not k_.startswith("radix_")}
    dom = max(single.items(), key=lambda kv: kv[1][1])[0]
    dom_ms = prof[dom][1] / prof[dom][0]
    dom_gbs = kb[dom] / (dom_ms * 1e-3) / 1e9 if kb.get(dom) else None
    traffic = measured_traffic(dom)

    # The frame's algorithmic bytes as the reference's two-stage algorithm
    # moves them (8-bit digits, 4 depth passes, ceil(tile_bits / 8) tile
    # passes), whatever this pipeline does instead.
    tile_bits = max(1, math.ceil(math.log2(isects.c.tiles_x * isects.c.tiles_y)))
    b_fwd_ref = (n * (44 + 4 * 3 * k) + 56 * vis + 16 * vis + 8 * ni + 4 * 16 * vis +
                 math.ceil(tile_bits / 8) * 16 * ni + 40 * ni + 16 * px)
    b_bwd_ref = 40 * ni + 32 * px + 36 * vis + 2 * (n * (44 + 4 * 3 * k)) + 36 * vis
    frame_gbs = (b_fwd_ref + b_bwd_ref) / (ms_per_step * 1e-3) / 1e9
    # The same for the reads and writes this pipeline must make.
    # (No placement pass: the rasterizer reads the sorted pairs instead.)
    b_fwd = (kb["project"] + depth_passes * 28 * n + 28 * pairs + kb["bin_scan"] +
             kb["bin_count"] + 8 * pairs + kb["raster_rgb"])
    b_bwd = 40 * ni + 32 * px + 36 * vis + n * (44 + 4 * 3 * k) + 36 * vis + n * (44 + 4 * 3 * k)
    step_gbs = (b_fwd + b_bwd) / (ms_per_step * 1e-3) / 1e9
    stage("copy ceiling: 256 MiB device copy")
    copy_gbs = copy_ceiling(ctx)

    extra = {}
    if not args.no_extra:
        extra["eig_round"] = bench_eig_round(ctx, args)
        stage("sort: 16M u64 keys")
        extra["sort_u64"] = bench_sort(ctx, args)
        extra["eig_round_c4"] = bench_c4(ctx, args)
        stage("paint: posterior-mean render at the frame's scene")
        extra["paint"] = bench_paint(ctx, splats, scratch, cam, w, h, n)
        stage("ply: decode the frame's scene")
        extra["ply_decode"] = bench_ply(ctx, attrs, sh, n)
        stage("sog: decode 1M SH3 splats")
        extra["sog_decode"] = bench_sog(ctx, n)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(attrs, sh, pa, ps, n, k, view, w, h, args.cpu_threads,
                           args.cpu_seconds)

    line = {
        "metric": "frames/s at 1M Gaussians 1080p fwd+bwd",
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
        "data": "synthetic (seeded splats: means U[-1,1]^3, log-scales N(-4,0.5), "
                "opacity logits N(0,1), SH3; target = same scene perturbed by N(0,0.01))",
        "config": {"workload": "1080p forward+backward frame, 1,048,576 Gaussians SH3, L2 loss",
                   "gaussians": n, "width": w, "height": h, "isects": ni, "visible": vis,
                   "supertile_pairs": pairs, "depth_sort_passes": depth_passes, "loss": loss_v},
        "roofline": {"bound": "hbm", "achieved": round(dom_gbs, 1) if dom_gbs else None,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(dom_gbs / HBM_PEAK_GBS, 4) if dom_gbs else None,
                     "traffic": traffic, "kernel": dom,
                     "frame_bytes": b_fwd_ref + b_bwd_ref,
                     "frame_achieved": round(frame_gbs, 1),
                     "frame_frac": round(frame_gbs / HBM_PEAK_GBS, 4),
                     "frame_bytes_implemented": b_fwd + b_bwd,
                     "frame_frac_implemented": round(step_gbs / HBM_PEAK_GBS, 4),
                     "copy_ceiling_gbs": copy_gbs,
                     "frame_frac_of_copy": round(frame_gbs / copy_gbs, 4)},
        "kernels": per_kernel,
        "extra": extra,
        "cpu_baseline": cpu,
        "wall_s": round(wall, 3),
    }
    print(json.dumps(line))
    ctx.close()


def measured_traffic(scope):
    """HBM bytes per launch of a kernel, from the rocprofv3 FETCH_SIZE and
    WRITE_SIZE summary committed under profiles/ (each counter scaled by
    what it reads for a device copy of known size), or None."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                        "traffic_fwd_bwd_1080p.json")
    try:
        with open(path) as f:
            t = json.load(f)
        return t["kernels"][scope]["hbm_bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        return None


def host_cores():
    """Cores this process may use, at most 16 (a GPU box's host share)."""
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    return max(1, min(16, avail))


def cpu_baseline(attrs, sh, pa, ps, n, k, view, w, h, threads, seconds):
    """The oracle's C restatement of the same fwd+bwd frame (f32 forward:
    project, sorts, map, composite; the CPU reference gradient of the same L2
    loss against the same perturbed-scene target), timed on the host: all
    cores for about `seconds`, then one thread for about the same.  The
    target render is outside the timed samples."""
    from oracle import oracle as O
    cores = threads or host_cores()
    ov = O.View.from_buffer_copy(bytes(view))
    target = O.render(ov, pa, ps, n, k, w, h, nthreads=cores)["rgba"]

    def sample(nthreads, budget):
        frames = 0
        t0 = time.perf_counter()
        while True:
            r = O.render(ov, attrs, sh, n, k, w, h, nthreads=nthreads)
            d = (2.0 * (r["rgba"] - target)).astype(np.float32)
            O.backward(ov, attrs, sh, n, k, w, h, r, r, d, nthreads=nthreads)
            frames += 1
            dt = time.perf_counter() - t0
            if dt >= budget:
                return frames, dt
    stage(f"cpu baseline: {cores} threads")
    fa, ta = sample(cores, seconds)
    stage("cpu baseline: 1 thread")
    f1, t1 = sample(1, seconds)
    return {"value": round(fa / ta, 4), "unit": "frames/s", "cores": cores, "kind": "port",
            "single_thread": round(f1 / t1, 4),
            "sample": f"the same 1080p fwd+bwd frame (1M SH3 splats, L2 loss against the "
                      f"perturbed-scene target) by the oracle's C restatement (f32 forward + "
                      f"CPU reference gradient): {fa} frames in {ta:.1f} s on {cores} threads; "
                      f"{f1} frames in {t1:.1f} s on 1 thread"}


def copy_ceiling(ctx, reps=5):
    """Measured device-copy bandwidth (256 MiB read + 256 MiB written per
    copy), GB/s: the practical ceiling under the 8 TB/s vendor peak."""
    from splatfield_amd.runtime import GpuTensor
    from splatfield_amd._lib import lib
    n = 64 << 20
    src = GpuTensor.empty(ctx, n)
    dst = GpuTensor.empty(ctx, n)
    src.fill_u32(7)
    ctx.check(lib().sfk_copy_u32(ctx.ptr, dst.ptr, src.ptr, n), "copy")
    ctx.sync()
    ctx.timer_start()
    for _ in range(reps):
        ctx.check(lib().sfk_copy_u32(ctx.ptr, dst.ptr, src.ptr, n), "copy")
    ms = ctx.timer_stop_ms() / reps
    return round(2 * n * 4 / (ms * 1e-3) / 1e9, 1)


if __name__ == "__main__":
    main()
