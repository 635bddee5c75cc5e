"""B3-Seg mirror (src/seg/mod.rs, beta.rs, views.rs) over libsfk:
fixed-point accumulators, mask packing, responsibility renders, analytic EIG
with f64 entropy on the device, and the Beta update."""
import ctypes as C
import math
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import lib
from .render import BASE_FOV, Camera
from .runtime import GpuTensor


def fractional_bits(w, h):
    """seg/mod.rs:25-34."""
    return int(lib().sfk_fractional_bits(w, h))


class Accumulators:
    """Fixed-point epsilon accumulators (seg/mod.rs:45-85)."""

    def __init__(self, ctx, total):
        self.bits = GpuTensor.empty(ctx, total)
        self.fg = GpuTensor.empty(ctx, total)
        self.bg = GpuTensor.empty(ctx, total)
        self.scale = 1.0

    def read_f32(self):
        return self.bits.read().astype(np.float32) / np.float32(self.scale)

    def read_evidence_f32(self):
        s = np.float32(self.scale)
        return (self.fg.read().astype(np.float32) / s, self.bg.read().astype(np.float32) / s)


class Mask:
    """Bit-packed 2D mask (seg/mod.rs:105-144)."""

    def __init__(self, ctx, w, h, bytes_):
        b = np.ascontiguousarray(bytes_, np.uint8)
        if b.size != w * h:
            raise ValueError("mask bytes must be one per pixel")
        self.words_per_row = (w + 31) // 32
        words = np.zeros(self.words_per_row * h, np.uint32)
        fg = C.c_uint32()
        lib().sfk_mask_pack(b.ctypes.data, w, h, words.ctypes.data, C.byref(fg))
        self.fg_pixels = fg.value
        self.words = words
        self.bits = GpuTensor.from_numpy(ctx, words)


def render_responsibility(splats, scratch, acc, camera, w, h):
    """Splats::render_responsibility (seg/mod.rs:146-159)."""
    ctx = splats.ctx
    c = camera.to_c()
    scale = C.c_float()
    ctx.check(lib().sfk_render_responsibility(ctx.ptr, splats.ptr, scratch.ptr, C.byref(c), w, h,
                                              acc.bits.ptr, C.byref(scale)),
              "render_responsibility")
    acc.scale = scale.value


def accumulate_prepared(scratch, isects, acc, mask=None):
    """accumulate_prepared (seg/mod.rs:161-210)."""
    ctx = scratch.ctx
    scale = C.c_float()
    if mask is None:
        rc = lib().sfk_accumulate_prepared(ctx.ptr, scratch.ptr, C.byref(isects.c), None, 0,
                                           acc.bits.ptr, None, None, C.byref(scale))
    else:
        rc = lib().sfk_accumulate_prepared(ctx.ptr, scratch.ptr, C.byref(isects.c),
                                           mask.bits.ptr, mask.words_per_row, None, acc.fg.ptr,
                                           acc.bg.ptr, C.byref(scale))
    ctx.check(rc, "accumulate_prepared")
    acc.scale = scale.value


class BetaState:
    """Per-splat Beta posteriors (beta.rs:214-287)."""

    def __init__(self, ctx, a, b):
        self.ctx = ctx
        self.a = GpuTensor.from_numpy(ctx, np.asarray(a, np.float32))
        self.b = GpuTensor.from_numpy(ctx, np.asarray(b, np.float32))
        self.n = self.a.shape[0]
        self._out = GpuTensor.empty(ctx, 1, np.float64)

    @classmethod
    def new_uniform(cls, ctx, total):
        return cls(ctx, np.ones(total, np.float32), np.ones(total, np.float32))

    def eig(self, acc):
        """Sum of per-splat entropy drops, f64 (SYNC)."""
        out = C.c_double()
        self.ctx.check(lib().sfk_eig(self.ctx.ptr, self.a.ptr, self.b.ptr, acc.bits.ptr, self.n,
                                     acc.scale, C.byref(out)), "eig")
        return out.value

    def update(self, acc):
        self.ctx.check(lib().sfk_beta_update(self.ctx.ptr, self.a.ptr, self.b.ptr, acc.fg.ptr,
                                             acc.bg.ptr, self.n, acc.scale), "beta update")

    def posteriors(self):
        return self.a.read(), self.b.read()


def map_labels(a, b):
    """MAP foreground iff a > b, ties background (beta.rs:289-297)."""
    a = np.asarray(a)
    b = np.asarray(b)
    if a.shape != b.shape:
        raise ValueError("posterior length mismatch")
    return a > b


def localize(splats, state):
    """views::localize (views.rs:32-67) on the device: (centre, radius) of
    the posterior-weighted foreground (a > b, weight a / (a + b)), or None
    when no splat is foreground."""
    ctx = splats.ctx
    center = np.zeros(3, np.float32)
    radius = C.c_float()
    found = C.c_int()
    ctx.check(lib().sfk_localize(ctx.ptr, splats.ptr, state.a.ptr, state.b.ptr,
                                 center.ctypes.data, C.byref(radius), C.byref(found)), "localize")
    if not found.value:
        return None
    return center, radius.value


CANDIDATE_FOV = BASE_FOV  # views.rs:69-71


def candidates(center, radius, count):
    """Fibonacci-sphere candidate cameras (views.rs:73-96)."""
    f = np.float32
    golden = f((math.sqrt(5.0) - 1.0) / 2.0)
    radius = max(f(radius), f(1e-3))
    dist = f(1.5) * radius / f(math.tan(f(CANDIDATE_FOV) * f(0.5)))
    cams = []
    c = np.asarray(center, np.float32)
    for k in range(count):
        z = f(1.0) - f(2.0) * (f(k) + f(0.5)) / f(count)
        r = f(math.sqrt(max(0.0, float(f(1.0) - z * z))))
        theta = f(2.0 * math.pi) * golden * f(k)
        d = np.array([r * f(math.cos(theta)), r * f(math.sin(theta)), z], np.float32)
        pos = c + dist * d
        cams.append(Camera.look_at(pos, c))
    return cams


class EigRound:
    """Batched EIG round over candidate cameras of mixed resolution
    (sfk_eig_round_*): one stream-ordered sequence, optionally replayed as a
    captured hipGraph."""

    def __init__(self, splats, cameras, sizes, tile_w=16, digit_bits=8):
        self.ctx = splats.ctx
        self.splats = splats
        self.nviews = len(cameras)
        cams = (_lib.Camera * self.nviews)(*[c.to_c() for c in cameras])
        self._w = np.ascontiguousarray([s[0] for s in sizes], np.uint32)
        self._h = np.ascontiguousarray([s[1] for s in sizes], np.uint32)
        out = C.c_void_p()
        self.ctx.check(lib().sfk_eig_round_create(self.ctx.ptr, splats.ptr, cams,
                                                  self._w.ctypes.data, self._h.ctypes.data,
                                                  self.nviews, tile_w, digit_bits, C.byref(out)),
                       "eig round create")
        self.ptr = out.value

    def set_binned(self, on):
        """Binned isect emission (default) or emission + stable tile argsort."""
        self.ctx.check(lib().sfk_eig_round_set_binned(self.ptr, 1 if on else 0), "binned")

    def set_pair_raster(self, on):
        """Binned only: raster straight from the supertile pairs (default) or
        from per-tile lists placed by the binned emission."""
        self.ctx.check(lib().sfk_eig_round_set_pair_raster(self.ptr, 1 if on else 0),
                       "pair raster")

    def run(self, state, use_graph=True, read=True):
        """Returns (eig per view, best view) when read is set."""
        eig = np.zeros(self.nviews, np.float64)
        best = C.c_uint32()
        self.ctx.check(lib().sfk_eig_round_run(self.ctx.ptr, self.ptr, state.a.ptr, state.b.ptr,
                                               int(use_graph),
                                               eig.ctypes.data if read else None,
                                               C.byref(best) if read else None), "eig round run")
        return (eig, best.value) if read else None

    def resp(self, view):
        """(u32 epsilon words of one view, its scale)."""
        p = C.c_void_p()
        sc = C.c_float()
        lib().sfk_eig_round_info(self.ptr, view, C.byref(p), C.byref(sc), None, None)
        t = GpuTensor(self.ctx, p.value, (self.splats.n,), np.uint32, False, self)
        return t.read(), sc.value

    def capacity(self):
        """Isect capacity of the round's buffers (refit on first graph use)."""
        cap = C.c_uint32()
        self.ctx.check(lib().sfk_eig_round_info(self.ptr, 0, None, None, C.byref(cap), None),
                       "eig round info")
        return cap.value

    def counters(self):
        out = np.zeros(2 * self.nviews + 1, np.uint32)
        self.ctx.check(lib().sfk_eig_round_counters(self.ctx.ptr, self.ptr, out.ctypes.data),
                       "eig round counters")
        return out

    def __del__(self):
        try:
            if self.ctx.ptr:  # the context must outlive the objects it made
                lib().sfk_eig_round_destroy(self.ptr)
        except Exception:
            pass


def eig_round_plain(splats, scratch, cameras, sizes, state, resp):
    """Segmenter::select_view (active.rs:113-129) in C: per candidate a
    responsibility render and a synchronous EIG; returns (eig, best)."""
    n = len(cameras)
    cams = (_lib.Camera * n)(*[c.to_c() for c in cameras])
    w = np.ascontiguousarray([s[0] for s in sizes], np.uint32)
    h = np.ascontiguousarray([s[1] for s in sizes], np.uint32)
    eig = np.zeros(n, np.float64)
    best = C.c_uint32()
    ctx = splats.ctx
    ctx.check(lib().sfk_eig_round_plain(ctx.ptr, splats.ptr, scratch.ptr, cams, w.ctypes.data,
                                        h.ctypes.data, n, state.a.ptr, state.b.ptr, resp.ptr,
                                        eig.ctypes.data, C.byref(best)), "plain eig round")
    return eig, best.value


def autotune_round(splats, cameras, sizes, state, reps=3):
    """Search tile width {8,16,32} x digit width {4,8} for the batched round;
    returns ((tile_w, digit_bits), {(tile_w, digit_bits): ms})."""
    n = len(cameras)
    cams = (_lib.Camera * n)(*[c.to_c() for c in cameras])
    w = np.ascontiguousarray([s[0] for s in sizes], np.uint32)
    h = np.ascontiguousarray([s[1] for s in sizes], np.uint32)
    tw, db = C.c_uint32(), C.c_uint32()
    ms = np.zeros(6, np.float32)
    ctx = splats.ctx
    ctx.check(lib().sfk_eig_round_autotune(ctx.ptr, splats.ptr, cams, w.ctypes.data,
                                           h.ctypes.data, n, state.a.ptr, state.b.ptr, reps,
                                           C.byref(tw), C.byref(db), ms.ctypes.data),
              "eig round autotune")
    keys = [(t, d) for t in (8, 16, 32) for d in (4, 8)]
    return (tw.value, db.value), dict(zip(keys, ms.tolist()))


def sphere_cameras(count, radius=3.0, center=(0.0, 0.0, 0.0), fov=CANDIDATE_FOV):
    """Fibonacci-sphere cameras at a fixed distance looking at the centre
    (views.rs:77-96 with dist fixed)."""
    from .scene import fibonacci_directions
    c = np.asarray(center, np.float32)
    dirs = fibonacci_directions(count)
    return [Camera.look_at(c + np.float32(radius) * d, c, fov=(fov, fov)) for d in dirs]


def footprint_mask(ctx, attrs, sh, n, sh_per_ch, camera, w, h, seed=0, radius=0.3,
                   threshold=0.5):
    """A synthetic 2D oracle's mask for one view: the rendered footprint of
    the splats within `radius` of a seeded centre (uniform in [-0.5, 0.5]^3).
    Only those splats are rendered; a pixel is foreground where their alpha
    1 - T reaches `threshold`.  Returns (one byte per pixel, the subset's
    boolean selection)."""
    from .render import RenderScratch, Splats, finalize_rgb, prepare_isects
    A = np.asarray(attrs, np.float32).reshape(-1, n)
    Sh = np.asarray(sh, np.float32).reshape(-1, n)
    centre = np.random.default_rng(seed).uniform(-0.5, 0.5, 3).astype(np.float32)
    sel = np.sqrt(((A[0:3].T - centre) ** 2).sum(axis=1)) <= np.float32(radius)
    m = int(sel.sum())
    if m == 0:
        return np.zeros(w * h, np.uint8), sel
    sub = Splats(ctx, np.ascontiguousarray(A[:, sel]).reshape(-1),
                 np.ascontiguousarray(Sh[:, sel]).reshape(-1))
    scratch = RenderScratch(ctx, m, w, h)
    isects = prepare_isects(sub, scratch, camera, w, h)
    finalize_rgb(scratch, isects)
    alpha = 1.0 - scratch.read("final_t").reshape(h, w)
    return (alpha >= threshold).astype(np.uint8).reshape(-1), sel


def observe(splats, scratch, acc, state, camera, w, h, mask_bytes):
    """The observation half of Segmenter::step (active.rs:150-176): render
    the view, lift the mask to fg/bg evidence over the render's own isects
    and fold it into the posterior.  Returns the mask's foreground pixels."""
    from .render import finalize_rgb, prepare_isects
    isects = prepare_isects(splats, scratch, camera, w, h)
    finalize_rgb(scratch, isects)
    mask = Mask(splats.ctx, w, h, mask_bytes)
    accumulate_prepared(scratch, isects, acc, mask)
    state.update(acc)
    return int(mask.fg_pixels)


def observe_bootstrap(splats, camera, res, mask_bytes):
    """A Beta(1,1) posterior after one observation of `mask_bytes` from
    `camera` at res x res (the round-0 step of the loop)."""
    from .render import RenderScratch
    n = splats.n
    state = BetaState.new_uniform(splats.ctx, n)
    observe(splats, RenderScratch(splats.ctx, n, res, res), Accumulators(splats.ctx, n), state,
            camera, res, res, mask_bytes)
    return state


def bitmap_to_rgb(bitmap_words, width, height, row_stride):
    """active.rs:240-256: the RGBA8 bitmap (row stride in words) unpacked to
    tightly packed RGB bytes (alpha dropped), row-major."""
    w = np.asarray(bitmap_words, np.uint32).reshape(height, row_stride)[:, :width]
    return np.ascontiguousarray(w.view(np.uint8).reshape(height, width, 4)[:, :, :3]).reshape(-1)


@dataclass(frozen=True)
class Config:
    """Loop parameters (active.rs:31-55): square render resolution, candidate
    views per iteration (N_cand) and active iterations (T), with the shipped
    defaults."""
    resolution: int = 512
    candidates: int = 20
    iterations: int = 20


@dataclass
class Iteration:
    """One oracle observation as the loop took it (active.rs:57-67): the
    camera, the EIG it scored (0 for the blind round-0 camera) and the mask's
    foreground pixels."""
    camera: Camera
    eig: float
    fg_pixels: int


class Segmenter:
    """Segmenter (active.rs:69-238) over the device pipeline.  Each step
    localizes the posterior foreground on the device, scores candidate views
    (batched round, or the plain select_view loop), renders the best view,
    hands its RGB bytes to the caller's 2D oracle, lifts the returned mask to
    per-splat evidence over the same intersections and updates the Beta
    posterior.  The config is validated before any device allocation
    (active.rs:86-95); errors are ValueError with the reference's messages."""

    def __init__(self, splats, cfg=Config(), camera=None, batched=True, tile_w=16):
        if cfg.iterations < 1:
            raise ValueError("iterations must be >= 1")
        if cfg.resolution < 16:
            raise ValueError("resolution must be >= 16")
        if cfg.candidates < 1:
            raise ValueError("candidates must be >= 1: zero would panic select_view's "
                             "cameras[0] index")
        from .render import RenderScratch
        self.cfg = cfg
        self.splats = splats
        self.ctx = splats.ctx
        self.camera = camera if camera is not None else Camera()
        self.batched = bool(batched)
        self.tile_w = tile_w
        n = splats.n
        res = cfg.resolution
        self.scratch = RenderScratch(self.ctx, n, res, res, tile_w=tile_w)
        self.acc = Accumulators(self.ctx, n)
        self.state = BetaState.new_uniform(self.ctx, n)
        self.bootstrapped = False

    def select_view(self, center, radius):
        """active.rs:110-129: the EIG-best of the candidate cameras (the first
        of equal scores, as the strict > of the reference's loop)."""
        cams = candidates(center, radius, self.cfg.candidates)
        sizes = [(self.cfg.resolution, self.cfg.resolution)] * len(cams)
        if self.batched:
            rnd = EigRound(self.splats, cams, sizes, tile_w=self.tile_w)
            eig, best = rnd.run(self.state, use_graph=False)
            del rnd
        else:
            resp = GpuTensor.empty(self.ctx, self.splats.n)
            eig, best = eig_round_plain(self.splats, self.scratch, cams, sizes, self.state, resp)
        return cams[best], float(eig[best])

    def step(self, oracle):
        """One observation round (active.rs:131-183).  oracle(rgb_bytes,
        (w, h), camera) returns one mask byte per pixel (nonzero =
        foreground); its exceptions surface as RuntimeError("oracle
        failed")."""
        from .render import finalize_rgb, prepare_isects
        loc = localize(self.splats, self.state)
        if loc is not None:
            camera, eig = self.select_view(*loc)
        elif not self.bootstrapped:
            camera, eig = self.camera, 0.0
        else:
            # nothing foreground yet: survey the whole scene
            lo, hi = (np.asarray(b, np.float32) for b in self.splats.bounds)
            camera, eig = self.select_view((lo + hi) * np.float32(0.5),
                                           float((hi - lo).max() * np.float32(0.5)))
        self.bootstrapped = True
        w = h = self.cfg.resolution
        isects = prepare_isects(self.splats, self.scratch, camera, w, h)
        finalize_rgb(self.scratch, isects)
        b = self.scratch.buffers()
        rgb = bitmap_to_rgb(self.scratch.read("bitmap"), w, h, b.row_stride)
        try:
            mask_bytes = oracle(rgb, (w, h), camera)
        except Exception as e:
            raise RuntimeError("oracle failed") from e
        mask = Mask(self.ctx, w, h, mask_bytes)
        accumulate_prepared(self.scratch, isects, self.acc, mask)
        self.state.update(self.acc)
        return Iteration(camera, float(eig), int(mask.fg_pixels))

    def run_with(self, oracle, on_round=None):
        """The loop (active.rs:185-225): up to cfg.iterations observations;
        on_round(iteration, state) returning False stops after the round just
        delivered; an empty mask ends the run while no round has seen
        foreground.  Returns the rounds taken."""
        any_fg = False
        rounds = []
        for _ in range(self.cfg.iterations):
            r = self.step(oracle)
            any_fg |= r.fg_pixels > 0
            stop = on_round is not None and not on_round(r, self.state)
            rounds.append(r)
            if stop:
                break
            if r.fg_pixels == 0 and not any_fg:
                break
        return rounds

    def posteriors(self):
        return self.state.posteriors()

    def segmentation(self):
        """The 3D segmentation so far, {i : a_i > b_i} (active.rs:227-232)."""
        return map_labels(*self.posteriors())
