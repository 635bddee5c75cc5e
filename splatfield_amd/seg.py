"""B3-Seg mirror (src/seg/mod.rs, beta.rs, views.rs) over libsfk:
fixed-point accumulators, mask packing, responsibility renders, analytic EIG
with f64 entropy on the device, and the Beta update."""
import ctypes as C
import math

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


def bitmap_to_rgb(bitmap_words, width, height, row_stride):
    """active.rs:240-256: the RGBA8 bitmap (row stride in words) unpacked to
    tightly packed RGB bytes (alpha dropped), row-major."""
    w = np.asarray(bitmap_words, np.uint32).reshape(height, row_stride)[:, :width]
    return np.ascontiguousarray(w.view(np.uint8).reshape(height, width, 4)[:, :, :3]).reshape(-1)


class Segmenter:
    """Segmenter (active.rs:60-183) over the device pipeline: each step
    localizes the posterior foreground on the device, scores candidate views
    (batched round or the plain select_view loop), renders the best view,
    hands its RGB bytes to the caller's 2D oracle, lifts the returned mask to
    per-splat evidence over the same intersections and updates the Beta
    posterior."""

    def __init__(self, splats, attributes, camera, resolution=512, num_candidates=20,
                 batched=True, tile_w=16):
        from .render import RenderScratch
        self.splats = splats
        self.ctx = splats.ctx
        self.camera = camera
        self.res = int(resolution)
        self.num_candidates = int(num_candidates)
        self.batched = bool(batched)
        self.tile_w = tile_w
        n = splats.n
        A = np.asarray(attributes, np.float32).reshape(-1)
        p = np.stack([A[0:n], A[n:2 * n], A[2 * n:3 * n]], 1)
        self.bounds = (p.min(0), p.max(0))  # render.rs:242-269
        self.scratch = RenderScratch(self.ctx, n, self.res, self.res, tile_w=tile_w)
        self.acc = Accumulators(self.ctx, n)
        self.state = BetaState.new_uniform(self.ctx, n)
        self.bootstrapped = False

    def select_view(self, center, radius):
        """active.rs:113-129: the EIG-best of the candidate cameras."""
        cams = candidates(center, radius, self.num_candidates)
        sizes = [(self.res, self.res)] * len(cams)
        if self.batched:
            rnd = EigRound(self.splats, cams, sizes, tile_w=self.tile_w)
            eig, best = rnd.run(self.state, use_graph=False)
            del rnd
        else:
            resp = GpuTensor.empty(self.ctx, self.splats.n)
            eig, best = eig_round_plain(self.splats, self.scratch, cams, sizes, self.state, resp)
        return cams[best], float(eig[best])

    def step(self, oracle):
        """active.rs:133-183.  oracle(rgb_bytes, (w, h), camera) -> one mask
        byte per pixel.  Returns (camera, eig, foreground pixels)."""
        from .render import finalize_rgb, prepare_isects
        loc = localize(self.splats, self.state)
        if loc is not None:
            camera, eig = self.select_view(*loc)
        elif not self.bootstrapped:
            camera, eig = self.camera, 0.0
        else:
            lo, hi = self.bounds
            camera, eig = self.select_view((lo + hi) * 0.5, float((hi - lo).max()) * 0.5)
        self.bootstrapped = True
        w = h = self.res
        isects = prepare_isects(self.splats, self.scratch, camera, w, h)
        finalize_rgb(self.scratch, isects)
        b = self.scratch.buffers()
        rgb = bitmap_to_rgb(self.scratch.read("bitmap"), w, h, b.row_stride)
        mask = Mask(self.ctx, w, h, oracle(rgb, (w, h), camera))
        accumulate_prepared(self.scratch, isects, self.acc, mask)
        self.state.update(self.acc)
        return camera, eig, mask.fg_pixels
