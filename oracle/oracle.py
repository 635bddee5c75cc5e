"""Oracle (test infrastructure only): numpy/ctypes front-end of the C
restatement in ``oracle/sfo.c``.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s
``cpu_baseline`` leg may import this module, and only as the checker.  The
product path (``splatfield_amd``) never imports it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "_build", "libsfo.so")

MODE_RGB, MODE_TOTAL, MODE_EVIDENCE = 0, 1, 2
PROJ_FLOATS = 9


class View(C.Structure):
    _fields_ = [
        ("rot0", C.c_float * 3), ("rot1", C.c_float * 3), ("rot2", C.c_float * 3),
        ("trans", C.c_float * 3), ("focal", C.c_float * 2), ("world_pos", C.c_float * 3),
        ("img", C.c_float * 2), ("tile_bounds", C.c_float * 2),
    ]

    def as_array(self):
        return np.frombuffer(bytes(self), dtype=np.float32).copy()


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _lib():
    if not os.path.exists(_SO):
        build()
    lib = C.CDLL(_SO)
    P = C.c_void_p
    u32, f32, f64, i32 = C.c_uint32, C.c_float, C.c_double, C.c_int
    sig = {
        "sfo_camera_view": (None, [P, P, P, u32, u32, u32, P]),
        "sfo_look_at": (None, [P, P, P]),
        "sfo_bits_for": (u32, [u32]),
        "sfo_fractional_bits": (u32, [u32, u32]),
        "sfo_project": (None, [P, P, P, u32, u32, u32, i32, P, P, P, P, P, P]),
        "sfo_depth_order": (u32, [P, P, u32, P]),
        "sfo_argsort_u32": (None, [P, P, u32, u32, P, P]),
        "sfo_argsort_u64": (None, [P, P, u32, u32, P, P]),
        "sfo_scan_gather": (None, [P, P, u32, P]),
        "sfo_map_isects": (None, [P, P, P, u32, u32, u32, P, P]),
        "sfo_lower_bound": (u32, [P, u32, u32]),
        "sfo_tile_ranges": (None, [P, u32, u32, P]),
        "sfo_rasterize": (None, [i32, u32, u32, u32, u32, f32, P, P, P, P, u32, i32,
                                 P, P, P, P, P, P, P]),
        "sfo_lgamma": (f64, [f64]),
        "sfo_digamma": (f64, [f64]),
        "sfo_beta_entropy": (f64, [f64, f64]),
        "sfo_eig": (f64, [P, P, P, u32, f32]),
        "sfo_beta_update": (None, [f32, P, P, u32, P, P]),
        "sfo_paint_posterior": (None, [P, P, P, u32]),
        "sfo_tint_colors": (None, [P, P, P, u32, f32, f32, f32]),
        "sfo_expf_export": (f32, [f32]),
        "sfo_logf_export": (f32, [f32]),
        "sfo_backward": (None, [P, P, P, u32, u32, u32, u32, u32, P, P, P, P, P, P, P, i32,
                                P, P, P, P, P, P, P, P]),
    }
    for name, (res, args) in sig.items():
        if not hasattr(lib, name):
            continue
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


LIB = _lib()


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


def expf(x):
    return np.float32(LIB.sfo_expf_export(float(x)))


def logf(x):
    return np.float32(LIB.sfo_logf_export(float(x)))


def camera_view(fov, pos, quat_xyzw, w, h, tile_w=16):
    v = View()
    fov = _f32(fov); pos = _f32(pos); q = _f32(quat_xyzw)
    LIB.sfo_camera_view(_p(fov), _p(pos), _p(q), w, h, tile_w, C.byref(v))
    return v


def look_at(pos, target):
    q = np.zeros(4, np.float32)
    p = _f32(pos); t = _f32(target)
    LIB.sfo_look_at(_p(p), _p(t), _p(q))
    return q


def bits_for(m):
    return int(LIB.sfo_bits_for(m))


def fractional_bits(w, h):
    return int(LIB.sfo_fractional_bits(w, h))


def project(view, attrs, sh, n, sh_per_ch, tile_w=16, nthreads=1):
    """project_splats (src/project.rs:234-388) -> dict of per-splat outputs."""
    attrs = _f32(attrs); sh = _f32(sh)
    out = dict(
        visible=np.zeros(n, np.uint8), depth_keys=np.zeros(n, np.uint32),
        projected=np.zeros(n * PROJ_FLOATS, np.float32), tile_counts=np.zeros(n, np.uint32),
        bbox=np.zeros(2 * n, np.uint32), counters=np.zeros(2, np.uint32))
    LIB.sfo_project(C.byref(view), _p(attrs), _p(sh), n, sh_per_ch, tile_w, nthreads,
                    _p(out["visible"]), _p(out["depth_keys"]), _p(out["projected"]),
                    _p(out["tile_counts"]), _p(out["bbox"]), _p(out["counters"]))
    return out


def depth_order(visible, keys):
    n = len(keys)
    order = np.zeros(max(n, 1), np.uint32)
    v = LIB.sfo_depth_order(_p(np.ascontiguousarray(visible, np.uint8)), _p(_u32(keys)), n,
                            _p(order))
    return order[:v].copy()


def argsort_u32(keys, vals, bits):
    keys = _u32(keys); vals = _u32(vals)
    ok = np.zeros_like(keys); ov = np.zeros_like(vals)
    LIB.sfo_argsort_u32(_p(keys), _p(vals), len(keys), bits, _p(ok), _p(ov))
    return ok, ov


def argsort_u64(keys, vals, bits):
    keys = np.ascontiguousarray(keys, np.uint64); vals = _u32(vals)
    ok = np.zeros_like(keys); ov = np.zeros_like(vals)
    LIB.sfo_argsort_u64(_p(keys), _p(vals), len(keys), bits, _p(ok), _p(ov))
    return ok, ov


def scan_gather(gids, counts):
    gids = _u32(gids); counts = _u32(counts)
    out = np.zeros(len(gids), np.uint32)
    LIB.sfo_scan_gather(_p(gids), _p(counts), len(gids), _p(out))
    return out


def map_isects(order, bbox, offsets, tiles_per_row, max_isects):
    order = _u32(order); bbox = _u32(bbox); offsets = _u32(offsets)
    tid = np.zeros(max(max_isects, 1), np.uint32); gid = np.zeros(max(max_isects, 1), np.uint32)
    LIB.sfo_map_isects(_p(order), _p(bbox), _p(offsets), tiles_per_row, max_isects, len(order),
                       _p(tid), _p(gid))
    return tid[:max_isects], gid[:max_isects]


def tile_ranges(tile_ids, num_tiles):
    tile_ids = _u32(tile_ids)
    r = np.zeros(2 * num_tiles, np.uint32)
    LIB.sfo_tile_ranges(_p(tile_ids), len(tile_ids), num_tiles, _p(r))
    return r


def lower_bound(ids, target):
    ids = _u32(ids)
    return int(LIB.sfo_lower_bound(_p(ids), len(ids), target))


def rasterize(mode, w, h, ranges, gaussian_ids, projected, n, tile_w=16, scale=0.0,
              mask=None, words_per_row=0, nthreads=1, row_stride=None):
    """rasterize_kernel (src/raster.rs:87-230) in one of its three modes."""
    if row_stride is None:
        row_stride = ((w * 4 + 255) // 256 * 256) // 4
    ranges = _u32(ranges); gaussian_ids = _u32(gaussian_ids); projected = _f32(projected)
    out = dict(final_t=np.zeros(w * h, np.float32), n_contrib=np.zeros(w * h, np.uint32))
    if mode == MODE_RGB:
        out["rgba"] = np.zeros(w * h * 4, np.float32)
        out["bitmap"] = np.zeros(h * row_stride, np.uint32)
    elif mode == MODE_TOTAL:
        out["resp"] = np.zeros(n, np.uint32)
    else:
        out["fg"] = np.zeros(n, np.uint32)
        out["bg"] = np.zeros(n, np.uint32)
    mask_a = None if mask is None else _u32(mask)
    LIB.sfo_rasterize(mode, w, h, tile_w, row_stride, scale, _p(ranges), _p(gaussian_ids),
                      _p(projected), _p(mask_a), words_per_row, nthreads,
                      _p(out.get("rgba")), _p(out.get("bitmap")), _p(out["final_t"]),
                      _p(out["n_contrib"]), _p(out.get("resp")), _p(out.get("fg")),
                      _p(out.get("bg")))
    return out


def prepare(view, attrs, sh, n, sh_per_ch, w, h, tile_w=16, max_isects=None, nthreads=1):
    """prepare_isects (src/render.rs:274-364): project, depth order, scan, map,
    stable tile sort, tile ranges."""
    pr = project(view, attrs, sh, n, sh_per_ch, tile_w, nthreads)
    order = depth_order(pr["visible"], pr["depth_keys"])
    offsets = scan_gather(order, pr["tile_counts"])
    raw = int(pr["counters"][0])
    cap = raw if max_isects is None else max_isects
    tx = (w + tile_w - 1) // tile_w
    ty = (h + tile_w - 1) // tile_w
    tid, gid = map_isects(order, pr["bbox"], offsets, tx, cap)
    ni = min(raw, cap)
    tid, gid = tid[:ni], gid[:ni]
    bits = bits_for(tx * ty)
    stid, sgid = argsort_u32(tid, gid, bits)
    ranges = tile_ranges(stid, tx * ty)
    pr.update(depth_order=order, offsets=offsets, tile_ids=stid, gaussian_ids=sgid,
              num_isects=ni, num_isects_raw=raw, num_visible=len(order), ranges=ranges,
              tiles_x=tx, tiles_y=ty, tile_bits=bits)
    return pr


def render(view, attrs, sh, n, sh_per_ch, w, h, tile_w=16, mode=MODE_RGB, scale=0.0,
           mask=None, words_per_row=0, nthreads=1):
    pr = prepare(view, attrs, sh, n, sh_per_ch, w, h, tile_w, nthreads=nthreads)
    ras = rasterize(mode, w, h, pr["ranges"], pr["gaussian_ids"], pr["projected"], n, tile_w,
                    scale, mask, words_per_row, nthreads)
    pr.update(ras)
    return pr


def lgamma(x):
    return LIB.sfo_lgamma(x)


def digamma(x):
    return LIB.sfo_digamma(x)


def beta_entropy(a, b):
    return LIB.sfo_beta_entropy(a, b)


def eig(a, b, resp, scale):
    a = _f32(a); b = _f32(b); resp = _u32(resp)
    return LIB.sfo_eig(_p(a), _p(b), _p(resp), len(a), scale)


def beta_update(a, b, fg, bg, scale):
    a = _f32(a).copy(); b = _f32(b).copy()
    LIB.sfo_beta_update(scale, _p(_u32(fg)), _p(_u32(bg)), len(a), _p(a), _p(b))
    return a, b


def parse_ply(data):
    """parse_ply (src/ply.rs:9-112) in numpy: header rules, then the rows
    transposed to field-major planes; the rotation normalised as glam's
    x86-64 Quat::normalize evaluates it ((x^2 + z^2) + (y^2 + w^2), sqrt,
    one f32 division per component).  Returns (attrs, sh) or raises
    ValueError with the reference's message."""
    data = bytes(data)
    pos, props, count, off = 0, [], 0, len(data)
    while pos < len(data):
        e = data.find(b"\n", pos)
        e = len(data) if e < 0 else e
        tok = data[pos:e].decode("latin-1").split()
        pos = min(e + 1, len(data))
        if not tok:
            continue
        if tok[0] == "end_header":
            off = pos
            break
        if tok[0] == "format" and len(tok) >= 2 and tok[1] != "binary_little_endian":
            raise ValueError(f"Unsupported PLY format: {tok[1]}")
        if len(tok) == 3 and tok[:2] == ["element", "vertex"]:
            count = int(tok[2])
        elif len(tok) == 3 and tok[0] == "property":
            if tok[1] != "float":
                raise ValueError(f"Unsupported property type {tok[1]}: {tok[2]}")
            props.append(tok[2])
    if count == 0:
        raise ValueError("PLY contains no vertices")
    names = ["x", "y", "z", "scale_0", "scale_1", "scale_2", "opacity", "rot_0", "rot_1",
             "rot_2", "rot_3", "f_dc_0", "f_dc_1", "f_dc_2"]
    col = {}
    for nm in names:
        if nm not in props:
            raise ValueError(f"Missing property: {nm}")
        col[nm] = props.index(nm)
    rest = sorted(((int(p[7:]), c) for c, p in enumerate(props)
                   if p.startswith("f_rest_") and p[7:].lstrip("+").isdigit()),
                  key=lambda t: t[0])
    stride = len(props)
    if (len(data) - off) // (4 * stride) < count:
        raise ValueError(f"failed to read vertex {(len(data) - off) // (4 * stride)} of {count}")
    rows = np.frombuffer(data, "<f4", count * stride, off).reshape(count, stride)
    f = np.float32
    x, y, z, w = (rows[:, col[k]] for k in ("rot_1", "rot_2", "rot_3", "rot_0"))
    ln = np.sqrt((x * x + z * z) + (y * y + w * w)).astype(f)
    attrs = np.stack([rows[:, col["x"]], rows[:, col["y"]], rows[:, col["z"]],
                      w / ln, x / ln, y / ln, z / ln, rows[:, col["scale_0"]],
                      rows[:, col["scale_1"]], rows[:, col["scale_2"]],
                      rows[:, col["opacity"]]]).astype(f)
    per = len(rest) // 3
    sh = np.zeros((len(rest) + 3, count), f)
    for c in range(3):
        sh[c] = rows[:, col[f"f_dc_{c}"]]
        for j in range(per):
            sh[(j + 1) * 3 + c] = rows[:, rest[c * per + j][1]]
    return attrs.reshape(-1), sh.reshape(-1)


def sog_unpack_quat(px, py, pz, tag):
    """unpack_quat (src/sog.rs:80-91) in f32: [w, x, y, z]."""
    f = np.float32
    u = lambda v: (f(v) / f(255.0) * f(2.0) - f(1.0)) / f(1.41421356)
    a, b, c = u(px), u(py), u(pz)
    d = np.sqrt(np.maximum(f(1.0) - a * a - b * b - c * c, f(0.0)))
    return {252: (d, a, b, c), 253: (a, d, b, c), 254: (a, b, d, c)}.get(int(tag), (a, b, c, d))


def sog_palette_offset(label, cw, cpc):
    """palette_offset (src/sog.rs:201-205): (column in texels, row)."""
    per_row = cw // cpc
    return label % per_row * cpc, label // per_row


def sog_decode(info, sheets, cw, codebooks):
    """The decode loop of parse_sog (src/sog.rs:108-195) in numpy over RGBA8
    sheets; exp_m1 and ln evaluated in f64 and rounded to f32 (the
    reference's f32 libm calls, correctly rounded but for rare ties).
    Returns (attrs, sh) or raises ValueError with the reference's message."""
    f = np.float32
    n, per = info["n"], info["sh_per_ch"]
    px = {k: v[:4 * n].reshape(n, 4) for k, v in sheets.items() if k != "centroids"}
    mins = np.asarray(info["mins"], f)
    spans = np.asarray(info["maxs"], f) - mins
    cb = codebooks.astype(f)
    attrs = np.zeros((11, n), f)
    for k in range(3):
        t = (px["lo"][:, k].astype(np.uint32) | px["hi"][:, k].astype(np.uint32) << 8)
        v = mins[k] + spans[k] * (t.astype(f) / f(65535.0))
        sgn = np.where(np.signbit(v), f(-1.0), f(1.0))
        attrs[k] = sgn * np.expm1(np.abs(v).astype(np.float64)).astype(f)
        attrs[7 + k] = cb[px["scales"][:, k]]
    tags = px["quats"][:, 3]
    if np.any(tags < 252):
        raise ValueError(f"quats: invalid rotation tag {tags[np.argmax(tags < 252)]}")
    for i in range(n):
        attrs[3:7, i] = sog_unpack_quat(*px["quats"][i])
    sh = np.zeros((3 * per, n), f)
    for k in range(3):
        sh[k] = cb[256 + px["sh0"][:, k].astype(np.int64)]
    e = np.clip(px["sh0"][:, 3].astype(f) / f(255.0), f(1e-6), f(1.0) - f(1e-6))
    attrs[10] = np.log((e / (f(1.0) - e)).astype(np.float64)).astype(f)
    if per > 1:
        cpc = per - 1
        cen = sheets["centroids"]
        pcount = cen.size // 4 // cpc
        for i in range(n):
            label = int(px["labels"][i, 0]) | int(px["labels"][i, 1]) << 8
            if label >= pcount:
                raise ValueError(f"shN label {label} >= palette size {pcount}")
            bx, by = sog_palette_offset(label, cw, cpc)
            for j in range(cpc):
                p = (by * cw + bx + j) * 4
                for k in range(3):
                    sh[(j + 1) * 3 + k, i] = cb[512 + int(cen[p + k])]
    return attrs.reshape(-1), sh.reshape(-1)


def write_ply(attrs, sh, n):
    """CpuSplats::write_ply (src/ply.rs:114-170): canonical INRIA property
    order, zero normals, f_rest channel-major."""
    attrs = _f32(attrs).reshape(11, n); sh = _f32(sh).reshape(-1, n)
    k = sh.shape[0] // 3
    rest = (k - 1) * 3
    props = ["x", "y", "z", "nx", "ny", "nz", "f_dc_0", "f_dc_1", "f_dc_2"]
    props += [f"f_rest_{j}" for j in range(rest)]
    props += ["opacity", "scale_0", "scale_1", "scale_2", "rot_0", "rot_1", "rot_2", "rot_3"]
    head = "ply\nformat binary_little_endian 1.0\nelement vertex %d\n" % n
    head += "".join(f"property float {p}\n" for p in props) + "end_header\n"
    cols = [attrs[0], attrs[1], attrs[2], np.zeros(n), np.zeros(n), np.zeros(n), sh[0], sh[1],
            sh[2]]
    per = rest // 3
    cols += [sh[(j % per + 1) * 3 + j // per] for j in range(rest)]
    cols += [attrs[10], attrs[7], attrs[8], attrs[9], attrs[3], attrs[4], attrs[5], attrs[6]]
    rows = np.stack(cols, axis=1).astype("<f4")
    return head.encode() + rows.tobytes()


def render_posterior(view, attrs, sh, n, sh_per_ch, w, h, a, b, tile_w=16, nthreads=1):
    """Splats::render_posterior (src/seg/paint.rs:22-47): prepare, paint each
    projected row with m = a/(a+b), RGB raster."""
    pr = prepare(view, attrs, sh, n, sh_per_ch, w, h, tile_w, nthreads=nthreads)
    a = _f32(a); b = _f32(b)
    proj = pr["projected"].copy()
    LIB.sfo_paint_posterior(_p(proj), _p(a), _p(b), n)
    ras = rasterize(MODE_RGB, w, h, pr["ranges"], pr["gaussian_ids"], proj, n, tile_w,
                    nthreads=nthreads)
    pr.update(ras)
    pr["projected"] = proj
    return pr


def tint(sh, a, b, color):
    """Splats::tint (src/seg/paint.rs:49-65) on a copy of the SH planes."""
    sh = _f32(sh).copy(); a = _f32(a); b = _f32(b)
    LIB.sfo_tint_colors(_p(sh), _p(a), _p(b), len(a), float(color[0]), float(color[1]),
                        float(color[2]))
    return sh


def localize(attrs, n, a, b):
    """views::localize (src/seg/views.rs:32-67) in f64: centre = sum(m p) /
    sum(m), radius = sum(m |p - centre|) / sum(m) over a > b, m = a/(a+b);
    None when no splat is foreground.  (The reference accumulates in f32 in
    index order; f64 is the value that loop approximates.)"""
    A = np.asarray(attrs, np.float32).reshape(-1)
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)
    fg = a > b
    if not fg.any():
        return None
    m = (a[fg] / (a[fg] + b[fg])).astype(np.float64)
    p = np.stack([A[0 * n:1 * n], A[1 * n:2 * n], A[2 * n:3 * n]], 1)[fg].astype(np.float64)
    sw = m.sum()
    c = (m[:, None] * p).sum(0) / sw
    r = (m * np.linalg.norm(p - c, axis=1)).sum() / sw
    return c, r


def tile_keys_u64(tile_ids, gaussian_ids, depth_keys):
    """The sorted u64 ``tile << 32 | depth`` keys of the two-stage sort."""
    return (np.asarray(tile_ids, np.uint64) << np.uint64(32)) | \
        np.asarray(depth_keys, np.uint64)[np.asarray(gaussian_ids, np.int64)]


# ------------------------------------------------------------- backward
_bsig = {
    "sfo_bwd_create": (C.c_void_p, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                    C.c_uint32, C.c_void_p, C.c_void_p, C.c_uint32]),
    "sfo_bwd_destroy": (None, [C.c_void_p]),
    "sfo_bwd_loss": (C.c_double, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int,
                                  C.c_void_p]),
    "sfo_bwd_grad": (None, [C.c_void_p] * 6),
}
for _n, (_r, _a) in _bsig.items():
    _f = getattr(LIB, _n)
    _f.restype = _r
    _f.argtypes = _a


class Loss64:
    """f64 L2 render loss with frozen branch decisions and its adjoint
    (oracle/sfo_bwd.c).  ``prep`` is an f32 ``prepare`` result that fixes the
    tile lists."""

    def __init__(self, view, n, sh_per_ch, w, h, prep, target, tile_w=16):
        self.n, self.k, self.w, self.h = n, sh_per_ch, w, h
        self._ranges = _u32(prep["ranges"])
        self._gids = _u32(prep["gaussian_ids"])
        self.target = _f32(target)
        self.ptr = LIB.sfo_bwd_create(C.byref(view), n, sh_per_ch, w, h, tile_w,
                                      _p(self._ranges), _p(self._gids), len(self._gids))

    def loss(self, attrs, sh, record=False, image=False):
        a = np.ascontiguousarray(attrs, np.float64)
        s = np.ascontiguousarray(sh, np.float64)
        img = np.zeros(self.w * self.h * 4) if image else None
        v = LIB.sfo_bwd_loss(self.ptr, _p(a), _p(s), _p(self.target), int(record), _p(img))
        return (v, img) if image else v

    def grad(self, attrs, sh):
        a = np.ascontiguousarray(attrs, np.float64)
        s = np.ascontiguousarray(sh, np.float64)
        self.loss(a, s, record=True)
        ga = np.zeros(self.n * 11)
        gs = np.zeros(self.n * 3 * self.k)
        LIB.sfo_bwd_grad(self.ptr, _p(a), _p(s), _p(self.target), _p(ga), _p(gs))
        return ga, gs

    def __del__(self):
        try:
            LIB.sfo_bwd_destroy(self.ptr)
        except Exception:
            pass


LIB.sfo_backward.restype = None
LIB.sfo_backward.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32,
                             C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p,
                             C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int,
                             C.c_void_p, C.c_void_p]


def backward(view, attrs, sh, n, sh_per_ch, w, h, prep, ras, d_rgba, tile_w=16, nthreads=1):
    """CPU reference gradient of the f32 render (oracle/sfo_bwd.c
    sfo_backward) from the f32 forward state ``prep`` (prepare) and ``ras``
    (rasterize RGB: final_t, n_contrib)."""
    ga = np.zeros(n * 11, np.float32)
    gs = np.zeros(n * 3 * sh_per_ch, np.float32)
    keep = [_f32(attrs), _f32(sh), _u32(prep["ranges"]), _u32(prep["gaussian_ids"]),
            _f32(prep["projected"]), _f32(ras["final_t"]), _u32(ras["n_contrib"]), _f32(d_rgba)]
    LIB.sfo_backward(C.byref(view), *[_p(a) for a in keep[:2]], n, sh_per_ch, w, h, tile_w,
                     *[_p(a) for a in keep[2:]], nthreads, _p(ga), _p(gs))
    return ga, gs
