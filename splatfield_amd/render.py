"""Render-path mirror (src/camera.rs, src/render.rs): Camera, Splats,
RenderScratch, prepare_isects / finalize / render_with over libsfk."""
import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import lib
from .runtime import GpuTensor
from .scene import ATTR_PLANES

BASE_FOV = 0.8  # camera.rs:25
PROJ_STRIDE = 12


@dataclass
class Camera:
    """Pinhole camera (camera.rs:4-19): per-axis fov (radians), position,
    rotation quaternion (x, y, z, w) mapping camera to world; the camera
    looks along its local +Z."""
    fov: tuple = (BASE_FOV, BASE_FOV)
    position: tuple = (0.0, 0.0, 0.0)
    rotation: tuple = (0.0, 0.0, 0.0, 1.0)

    def to_c(self):
        c = _lib.Camera()
        c.fov[:] = [float(v) for v in self.fov]
        c.position[:] = [float(v) for v in self.position]
        c.rotation[:] = [float(v) for v in self.rotation]
        return c

    def focal(self, w, h):
        """camera.rs:38-40 (f32)."""
        f = np.float32
        return (f(f(w) * f(0.5)) / f(math.tan(f(self.fov[0]) * f(0.5))),
                f(f(h) * f(0.5)) / f(math.tan(f(self.fov[1]) * f(0.5))))

    @staticmethod
    def look_at(position, target, fov=(BASE_FOV, BASE_FOV)):
        """Camera::look_at (camera.rs:57-63)."""
        p = np.ascontiguousarray(position, np.float32)
        t = np.ascontiguousarray(target, np.float32)
        q = np.zeros(4, np.float32)
        lib().sfk_look_at(p.ctypes.data, t.ctypes.data, q.ctypes.data)
        return Camera(tuple(fov), tuple(p.tolist()), tuple(q.tolist()))

    def forward(self):
        """The viewing direction, rotation * +Z (f64)."""
        x, y, z, w = (float(v) for v in self.rotation)
        return np.array([2.0 * (x * z + w * y), 2.0 * (y * z - w * x), 1.0 - 2.0 * (x * x + y * y)])

    def frame_bounds(self, bounds):
        """Camera::frame_bounds (camera.rs:46-52): centre on the bounds and
        pull back along -Y by twice their largest extent, pitched -90 degrees
        about X (looking along +Y).  Returns that distance."""
        f = np.float32
        lo, hi = (np.asarray(b, np.float32) for b in bounds)
        d = f((hi - lo).max() * f(2.0))
        self.position = tuple(((lo + hi) * f(0.5) - np.array([0, d, 0], np.float32)).tolist())
        half = f(-math.pi / 2) * f(0.5)
        self.rotation = (float(f(math.sin(half))), 0.0, 0.0, float(f(math.cos(half))))
        return float(d)

    def view(self, w, h, tile_w=16):
        """CameraViewLaunch::for_camera (project.rs:76-96)."""
        v = _lib.View()
        c = self.to_c()
        rc = lib().sfk_camera_view(C.byref(c), w, h, tile_w, C.byref(v))
        if rc != 0:
            raise ValueError(f"camera view rejected ({rc})")
        return v


def parse_ply_header(data):
    """sfk_ply_parse_header (src/ply.rs:13-75): the header's layout, or
    ValueError with the reference's message."""
    buf = np.frombuffer(bytes(data), np.uint8) if not isinstance(data, np.ndarray) else data
    lay = _lib.PlyLayout()
    err = C.create_string_buffer(256)
    if lib().sfk_ply_parse_header(buf.ctypes.data, buf.size, C.byref(lay), err, 256) != 0:
        raise ValueError(err.value.decode())
    return lay


def view_array(v):
    return np.frombuffer(bytes(v), dtype=np.float32).copy()


class Splats:
    """Uploaded field-major splats (render.rs:200-269)."""

    def __init__(self, ctx, attributes, sh_coeffs):
        attributes = np.ascontiguousarray(attributes, np.float32)
        sh_coeffs = np.ascontiguousarray(sh_coeffs, np.float32)
        n = attributes.size // ATTR_PLANES
        if n == 0:
            raise ValueError("Splats::new: zero splats")
        per_ch = sh_coeffs.size // n // 3
        self.ctx = ctx
        out = C.c_void_p()
        ctx.check(lib().sfk_splats_create(ctx.ptr, attributes.ctypes.data, sh_coeffs.ctypes.data,
                                          n, per_ch, C.byref(out)), "Splats::new")
        self.ptr = out.value
        self.n = n
        self.sh_per_ch = per_ch
        xyz = attributes[:3 * n].reshape(3, n)
        self.bounds = (xyz.min(axis=1), xyz.max(axis=1))
        self.positions = xyz.copy()

    @classmethod
    def from_ply(cls, ctx, data):
        """parse_ply (src/ply.rs:9-112) + Splats::new: header parsed on the
        host, vertex rows decoded into the field-major planes on the device.
        data: the file's bytes.  Errors carry the reference's messages."""
        buf = np.frombuffer(bytes(data), np.uint8)
        out = C.c_void_p()
        ctx.check(lib().sfk_splats_from_ply(ctx.ptr, buf.ctypes.data, buf.size, C.byref(out)),
                  "parse_ply")
        self = cls.__new__(cls)
        self.ctx, self.ptr = ctx, out.value
        lay = parse_ply_header(buf)
        self.n, self.sh_per_ch = int(lay.n), int(lay.sh_per_ch)
        d_attrs, _ = self.device_ptrs()
        xyz = GpuTensor(ctx, d_attrs, (3, self.n), np.float32, False, self).read()
        self.bounds = (xyz.min(axis=1), xyz.max(axis=1))
        self.positions = xyz
        return self

    @classmethod
    def from_sog(cls, ctx, data):
        """parse_sog (src/sog.rs:97-198) + Splats::new: sheets unzipped and
        WebP-decoded on the host, every splat decoded on the device."""
        from . import scene as S
        info, sheets, cw, cbs = S.read_sog(data)
        n, per = info["n"], info["sh_per_ch"]
        m = _lib.SogMeta()
        m.n = n
        m.mins[:] = info["mins"]
        m.maxs[:] = info["maxs"]
        m.sh_per_ch = per
        if per > 1:
            m.centroid_w = cw
            m.centroid_px = sheets["centroids"].size // 4
            m.palette_count = m.centroid_px // (per - 1)
        dev = {k: GpuTensor.from_numpy(ctx, v) for k, v in sheets.items()}
        d_cb = GpuTensor.from_numpy(ctx, cbs)
        d_attrs = GpuTensor.empty(ctx, 11 * n, np.float32)
        d_sh = GpuTensor.empty(ctx, 3 * per * n, np.float32)
        status = GpuTensor.empty(ctx, 2, np.uint32)
        ptr = lambda k: dev[k].ptr if k in dev else None
        ctx.check(lib().sfk_sog_decode(ctx.ptr, C.byref(m), ptr("lo"), ptr("hi"), ptr("scales"),
                                       ptr("quats"), ptr("sh0"), ptr("labels"),
                                       ptr("centroids"), d_cb.ptr, d_attrs.ptr, d_sh.ptr,
                                       status.ptr), "parse_sog")
        bad_tag, bad_label = (int(v) for v in status.read())
        if bad_tag != 0xFFFFFFFF:
            raise ValueError(f"quats: invalid rotation tag {sheets['quats'][4 * bad_tag + 3]}")
        if bad_label != 0xFFFFFFFF:
            lb = sheets["labels"]
            label = int(lb[4 * bad_label]) | int(lb[4 * bad_label + 1]) << 8
            raise ValueError(f"shN label {label} >= palette size {m.palette_count}")
        out = C.c_void_p()
        ctx.check(lib().sfk_splats_create_device(ctx.ptr, d_attrs.ptr, d_sh.ptr, n, per,
                                                 C.byref(out)), "Splats::new")
        self = cls.__new__(cls)
        self.ctx, self.ptr, self.n, self.sh_per_ch = ctx, out.value, n, per
        xyz = d_attrs.read()[:3 * n].reshape(3, n)
        self.bounds = (xyz.min(axis=1), xyz.max(axis=1))
        self.positions = xyz
        return self

    def render_posterior(self, scratch, a, b, camera, w, h):
        """Splats::render_posterior (src/seg/paint.rs:22-47): the frame with
        every splat painted by its posterior mean a/(a+b) as grayscale; a
        posterior of another length renders the normal colours.  a, b are
        device GpuTensors; returns the RGBA8 bitmap (h x row_stride u32)."""
        ctx = self.ctx
        c = camera.to_c()
        bm = C.c_void_p()
        n_ab = a.shape[0] if a.shape == b.shape else 0
        ctx.check(lib().sfk_render_posterior(ctx.ptr, self.ptr, scratch.ptr, a.ptr, b.ptr, n_ab,
                                             C.byref(c), w, h, C.byref(bm), None),
                  "render_posterior")
        bf = scratch.buffers()
        return GpuTensor(ctx, bm.value, (h, bf.row_stride), np.uint32, False, scratch).read()

    def tint(self, a, b, color):
        """Splats::tint (src/seg/paint.rs:49-65): DC planes of the splats
        with a > b set to color, in place on the device."""
        if a.shape != b.shape:
            raise ValueError(f"tint: a {a.shape} vs b {b.shape}")
        col = np.ascontiguousarray(color, np.float32)
        self.ctx.check(lib().sfk_tint(self.ctx.ptr, self.ptr, a.ptr, b.ptr, a.shape[0],
                                      col.ctypes.data), "tint")

    def save_colors(self):
        """Splats::save_colors (paint.rs:67-74): device snapshot [3, n] of
        the DC planes."""
        saved = GpuTensor.empty(self.ctx, (3, self.n), np.float32)
        self.ctx.check(lib().sfk_save_colors(self.ctx.ptr, self.ptr, saved.ptr), "save_colors")
        return saved

    def restore_colors(self, saved):
        """Splats::restore_colors (paint.rs:76-79)."""
        if saved.shape != (3, self.n):
            raise ValueError(f"restore_colors: snapshot {saved.shape} for {self.n} splats")
        self.ctx.check(lib().sfk_restore_colors(self.ctx.ptr, self.ptr, saved.ptr),
                       "restore_colors")

    def device_ptrs(self):
        a, s = C.c_void_p(), C.c_void_p()
        lib().sfk_splats_info(self.ptr, None, None, None, None, C.byref(a), C.byref(s))
        return a.value, s.value

    def __del__(self):
        try:
            if self.ctx.ptr:  # the context must outlive the objects it made
                lib().sfk_splats_destroy(self.ptr)
        except Exception:
            pass


class TileIsects:
    """Sorted per-tile intersections (render.rs:463-471); aliases the scratch."""

    def __init__(self, scratch, c):
        self.scratch = scratch
        self.c = c

    @property
    def tile_bounds(self):
        return (self.c.tiles_x, self.c.tiles_y)

    def counts(self):
        """SYNC: (isects_raw, visible, isects)."""
        out = np.zeros(3, np.uint32)
        ctx = self.scratch.ctx
        ctx.check(lib().sfk_tile_isects_counts(ctx.ptr, C.byref(self.c), out.ctypes.data),
                  "isect counts")
        return int(out[0]), int(out[1]), int(out[2])

    def tile_ids(self, n):
        if not self.c.tile_ids:
            raise ValueError("tile ids were not written (RenderScratch.set_tile_ids(False))")
        return GpuTensor(self.scratch.ctx, self.c.tile_ids, (n,), np.uint32, False,
                         self.scratch).read()

    def gaussian_ids(self, n):
        return GpuTensor(self.scratch.ctx, self.c.gaussian_ids, (n,), np.uint32, False,
                         self.scratch).read()

    def tile_starts(self):
        t = self.c.tiles_x * self.c.tiles_y
        return GpuTensor(self.scratch.ctx, self.c.tile_starts, (t + 1,), np.uint32, False,
                         self.scratch).read()


class RenderScratch:
    """Per-frame buffers reused across frames (render.rs:114-198)."""

    def __init__(self, ctx, total, w, h, tile_w=16, digit_bits=8):
        self.ctx = ctx
        out = C.c_void_p()
        ctx.check(lib().sfk_render_scratch_create(ctx.ptr, total, w, h, C.byref(out)),
                  "RenderScratch::new")
        self.ptr = out.value
        if tile_w != 16:
            self.set_tile_width(tile_w)
        if digit_bits != 8:
            self.set_digit_bits(digit_bits)

    def set_tile_width(self, tw):
        self.ctx.check(lib().sfk_render_scratch_set_tile_width(self.ptr, tw), "tile width")

    def set_digit_bits(self, bits):
        self.ctx.check(lib().sfk_render_scratch_set_digit_bits(self.ptr, bits), "digit bits")

    def set_binned(self, on):
        """Binned isect emission (default) or emission + stable tile argsort;
        both give the same isect arrays on grids of at most 255 x 255 tiles."""
        self.ctx.check(lib().sfk_render_scratch_set_binned(self.ptr, 1 if on else 0), "binned")

    def set_tile_ids(self, on):
        """Whether the binned emission writes per-isect tile ids (default on);
        off, TileIsects.tile_ids is NULL and the rasterizers run unchanged."""
        self.ctx.check(lib().sfk_render_scratch_set_tile_ids(self.ptr, 1 if on else 0),
                       "tile ids")

    def set_pair_raster(self, on):
        """Pair raster (default off): the rasterizers take each tile's isects
        from its supertile's sorted pairs and write back only the ones they
        consume; images, evidence and gradients are unchanged."""
        self.ctx.check(lib().sfk_render_scratch_set_pair_raster(self.ptr, 1 if on else 0),
                       "pair raster")

    def reserve_isects(self, n):
        self.ctx.check(lib().sfk_render_scratch_reserve_isects(self.ptr, n), "reserve")

    def buffers(self):
        b = _lib.RenderBuffers()
        self.ctx.check(lib().sfk_render_scratch_buffers(self.ptr, C.byref(b)), "buffers")
        return b

    def read(self, name):
        """Download one named scratch buffer as numpy."""
        b = self.buffers()
        n, w, h = b.n, b.w, b.h
        spec = {
            "depth_keys": (n, np.uint32), "depth_order": (n, np.uint32),
            "projected": (n * PROJ_STRIDE, np.float32), "tile_counts": (n, np.uint32),
            "bbox": (2 * n, np.uint32), "offsets": (n, np.uint32), "counters": (2, np.uint32),
            "bitmap": (h * b.row_stride, np.uint32), "rgba": (w * h * 4, np.float32),
            "final_t": (w * h, np.float32), "n_contrib": (w * h, np.uint32),
        }
        count, dt = spec[name]
        return GpuTensor(self.ctx, getattr(b, name), (count,), dt, False, self).read()

    def device(self, name):
        b = self.buffers()
        return getattr(b, name)

    def __del__(self):
        try:
            if self.ctx.ptr:  # the context must outlive the objects it made
                lib().sfk_render_scratch_destroy(self.ptr)
        except Exception:
            pass


def prepare_isects(splats, scratch, camera=None, w=None, h=None, view=None):
    """Splats::prepare_isects (render.rs:271-364)."""
    out = _lib.TileIsects()
    ctx = splats.ctx
    if view is not None:
        ctx.check(lib().sfk_prepare_isects_view(ctx.ptr, splats.ptr, scratch.ptr, C.byref(view),
                                                w, h, C.byref(out)), "prepare_isects")
    else:
        c = camera.to_c()
        ctx.check(lib().sfk_prepare_isects(ctx.ptr, splats.ptr, scratch.ptr, C.byref(c), w, h,
                                           C.byref(out)), "prepare_isects")
    return TileIsects(scratch, out)


def finalize_rgb(scratch, isects):
    ctx = scratch.ctx
    ctx.check(lib().sfk_finalize_rgb(ctx.ptr, scratch.ptr, C.byref(isects.c)), "finalize(Rgb)")


def finalize_total(scratch, isects, scale, resp):
    ctx = scratch.ctx
    ctx.check(lib().sfk_finalize_total(ctx.ptr, scratch.ptr, C.byref(isects.c), scale, resp.ptr),
              "finalize(Total)")


def finalize_evidence(scratch, isects, scale, mask, words_per_row, fg, bg):
    ctx = scratch.ctx
    ctx.check(lib().sfk_finalize_evidence(ctx.ptr, scratch.ptr, C.byref(isects.c), scale,
                                          mask.ptr, words_per_row, fg.ptr, bg.ptr),
              "finalize(Evidence)")


def render_with(splats, scratch, camera, w, h):
    """Splats::render_with (render.rs:366-380): returns the RGBA8 bitmap
    (h x row_stride u32) as numpy."""
    ctx = splats.ctx
    c = camera.to_c()
    bm, rgba = C.c_void_p(), C.c_void_p()
    ctx.check(lib().sfk_render_with(ctx.ptr, splats.ptr, scratch.ptr, C.byref(c), w, h,
                                    C.byref(bm), C.byref(rgba)), "render_with")
    b = scratch.buffers()
    return GpuTensor(ctx, bm.value, (h, b.row_stride), np.uint32, False, scratch).read()


def loss_l2(ctx, rgba, target, pixels, d_rgba, d_loss=None):
    """d_rgba = 2 (rgba - target); d_loss (device f64) = sum of squares."""
    ctx.check(lib().sfk_loss_l2(ctx.ptr, rgba, target, pixels, d_rgba,
                                d_loss.ptr if d_loss is not None else None), "loss_l2")


def render_backward_l2(splats, scratch, isects, rgba, target, grad_attrs, grad_sh, d_loss):
    """loss_l2 + render_backward fused (sfk_render_backward_l2): the adjoint
    forms d_rgba = 2 (rgba - target) itself; d_loss (device f64) receives
    the loss."""
    ctx = splats.ctx
    ctx.check(lib().sfk_render_backward_l2(ctx.ptr, splats.ptr, scratch.ptr, C.byref(isects.c),
                                           rgba, target, grad_attrs.ptr, grad_sh.ptr, d_loss.ptr),
              "render_backward_l2")


def render_backward(splats, scratch, isects, d_rgba, grad_attrs, grad_sh):
    """Adjoint of finalize(Rgb) + project_splats (sfk_render_backward)."""
    ctx = splats.ctx
    ctx.check(lib().sfk_render_backward(ctx.ptr, splats.ptr, scratch.ptr, C.byref(isects.c),
                                        d_rgba, grad_attrs.ptr, grad_sh.ptr), "render_backward")
