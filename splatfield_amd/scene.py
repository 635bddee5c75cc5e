"""Synthetic splat scenes and the camera set-ups of the benchmark workloads.

Layout contract (src/layout.rs:44-61): attribute plane ``k`` of splat ``i``
lives at ``attrs[k * n + i]`` in the order x, y, z, qw, qx, qy, qz, sx, sy,
sz, opacity (log-scales, opacity logit, normalised quaternion).  SH
coefficient ``k`` of channel ``c`` lives at ``sh[(k * 3 + c) * n + i]``.
"""
import math

import numpy as np

ATTR_PLANES = 11
PLANE_X, PLANE_Y, PLANE_Z = 0, 1, 2
PLANE_QW, PLANE_QX, PLANE_QY, PLANE_QZ = 3, 4, 5, 6
PLANE_SX, PLANE_SY, PLANE_SZ = 7, 8, 9
PLANE_OPACITY = 10
PROJ_FLOATS = 9


def synthetic_scene(n, sh_degree, seed=0):
    """Seeded synthetic splats: means U[-1,1]^3, log-scales N(-4, 0.5),
    quaternions normalize(N(0,1)^4) (stored normalised, as the PLY loader
    does, src/ply.rs:84-85), opacity logits N(0,1), SH band 0 N(0, 0.3) and
    higher bands N(0, 0.05).  Draw order: means, log-scales, quaternions,
    opacities, SH DC, SH rest.  Returns (attrs[11*n], sh[3*K*n], K)."""
    rng = np.random.default_rng(seed)
    k = (sh_degree + 1) ** 2
    attrs = np.empty((ATTR_PLANES, n), np.float32)
    attrs[0:3] = rng.uniform(-1.0, 1.0, size=(3, n)).astype(np.float32)
    log_scales = rng.normal(-4.0, 0.5, size=(3, n)).astype(np.float32)
    q = rng.normal(0.0, 1.0, size=(4, n)).astype(np.float32)
    q /= np.sqrt((q * q).sum(axis=0, keepdims=True))
    attrs[3:7] = q
    attrs[7:10] = log_scales
    attrs[10] = rng.normal(0.0, 1.0, size=n).astype(np.float32)
    sh = np.empty((k * 3, n), np.float32)
    sh[0:3] = rng.normal(0.0, 0.3, size=(3, n)).astype(np.float32)
    if k > 1:
        sh[3:] = rng.normal(0.0, 0.05, size=((k - 1) * 3, n)).astype(np.float32)
    return attrs.reshape(-1), sh.reshape(-1), k


def fov_for_aspect(w, h, fov_y):
    """Horizontal fov that keeps square pixels for a vertical fov
    (the aspect rule of src/camera.rs:27-36 applied to a fixed fov_y)."""
    return 2.0 * math.atan((w / h) * math.tan(fov_y * 0.5))


def fibonacci_directions(count):
    """Unit directions on the Fibonacci sphere (src/seg/views.rs:81-91)."""
    golden = np.float32((math.sqrt(5.0) - 1.0) / 2.0)
    out = []
    for k in range(count):
        z = np.float32(1.0) - np.float32(2.0) * (np.float32(k) + np.float32(0.5)) / np.float32(count)
        r = np.float32(math.sqrt(max(0.0, float(np.float32(1.0) - z * z))))
        theta = np.float32(2.0 * math.pi) * golden * np.float32(k)
        out.append((np.float32(r * np.float32(math.cos(theta))),
                    np.float32(r * np.float32(math.sin(theta))), np.float32(z)))
    return np.array(out, np.float32)


def write_ply(attrs, sh, n):
    """CpuSplats::write_ply (src/ply.rs:114-170): binary little-endian 3DGS
    PLY with the canonical INRIA property order (x y z, zero normals, f_dc,
    f_rest channel-major, opacity, scales, rot_0 = w then x y z), the exact
    inverse of the device parse for normalised rotations."""
    attrs = np.ascontiguousarray(attrs, np.float32).reshape(11, n)
    sh = np.ascontiguousarray(sh, np.float32).reshape(-1, n)
    if n == 0 or sh.shape[0] % 3:
        raise ValueError("no splats to write")
    per = sh.shape[0] // 3 - 1
    names = ["x", "y", "z", "nx", "ny", "nz", "f_dc_0", "f_dc_1", "f_dc_2"]
    names += ["f_rest_%d" % j for j in range(3 * per)]
    names += ["opacity", "scale_0", "scale_1", "scale_2", "rot_0", "rot_1", "rot_2", "rot_3"]
    header = ("ply\nformat binary_little_endian 1.0\nelement vertex %d\n" % n +
              "".join("property float %s\n" % p for p in names) + "end_header\n")
    out = np.zeros((n, len(names)), "<f4")
    out[:, 0:3] = attrs[0:3].T
    out[:, 6:9] = sh[0:3].T
    for c in range(3):  # file column 9 + c*per + j holds coefficient j+1, channel c
        out[:, 9 + c * per:9 + (c + 1) * per] = sh[3 + c::3][:per].T
    t = 9 + 3 * per
    out[:, t] = attrs[10]
    out[:, t + 1:t + 4] = attrs[7:10].T
    out[:, t + 4:t + 8] = attrs[3:7].T
    return header.encode() + out.tobytes()


def _sog_meta_schema(meta):
    """The shape serde demands of meta.json (sog.rs:8-42): count a
    non-negative integer; means {mins, maxs: 3 numbers, files: 2 names};
    scales and sh0 {codebook: numbers, files: names}; quats {files: 1 name};
    shN absent/null or {bands: non-negative integer, codebook: numbers,
    files: 2 names}.  Returns (count, files by sheet, shN or None)."""
    def uint(v):
        if isinstance(v, bool) or not isinstance(v, int) or v < 0:
            raise TypeError("not a non-negative integer")
        return v

    def nums(v, k=None):
        if not isinstance(v, list) or (k is not None and len(v) != k) or \
                not all(isinstance(x, (int, float)) and not isinstance(x, bool) for x in v):
            raise TypeError("not a number array")
        return v

    def names(v, k=None):
        if not isinstance(v, list) or (k is not None and len(v) != k) or not v or \
                not all(isinstance(x, str) for x in v):
            raise TypeError("not a file-name array")
        return v

    if not isinstance(meta, dict):
        raise TypeError("not an object")
    n = uint(meta["count"])
    means, scales, quats, sh0 = meta["means"], meta["scales"], meta["quats"], meta["sh0"]
    nums(means["mins"], 3)
    nums(means["maxs"], 3)
    mf = names(means["files"], 2)
    nums(scales["codebook"])
    sf = names(scales["files"])
    qf = names(quats["files"], 1)
    nums(sh0["codebook"])
    cf = names(sh0["files"])
    files = dict(lo=mf[0], hi=mf[1], scales=sf[0], quats=qf[0], sh0=cf[0])
    shn = meta.get("shN")
    if shn is not None:
        uint(shn["bands"])
        nums(shn["codebook"])
        nf = names(shn["files"], 2)
        files.update(centroids=nf[0], labels=nf[1])
    return n, files, shn


def read_sog(data):
    """The host half of parse_sog (src/sog.rs:44-64, 97-198): unzip, read
    meta.json, decode the WebP sheets to RGBA8, with the reference's checks
    and messages in its order (meta schema, count, means / scales sheets,
    quats sheet and rotation tags, shN bands, sh0, shN sheets and labels).
    Returns (meta dict, sheets dict of flat uint8 RGBA arrays, centroid
    sheet width, codebooks float32[768]).  A label inside the palette count
    whose palette runs off the centroid sheet (the reference indexes past
    the sheet there) is refused with its own message."""
    import io
    import json
    import zipfile
    try:
        z = zipfile.ZipFile(io.BytesIO(bytes(data)))
    except zipfile.BadZipFile:
        raise ValueError("not a SOG archive") from None
    names = set(z.namelist())
    if "meta.json" not in names:
        raise ValueError("missing meta.json")
    try:
        meta = json.loads(z.read("meta.json"))
        n, files, shn = _sog_meta_schema(meta)
    except (ValueError, KeyError, TypeError, IndexError):
        raise ValueError("invalid meta.json") from None
    if n == 0:
        raise ValueError("SOG contains no splats")
    means, scales, sh0 = meta["means"], meta["scales"], meta["sh0"]

    def rgba(name, min_px):
        if name not in names:
            raise ValueError(f"missing {name}")
        from PIL import Image
        try:
            img = Image.open(io.BytesIO(z.read(name)), formats=["WEBP"]).convert("RGBA")
        except Exception:
            raise ValueError(f"decode {name}") from None
        w, h = img.size
        if w * h < min_px:
            raise ValueError(f"{name}: {w}x{h} < {min_px} pixels")
        return np.frombuffer(img.tobytes(), np.uint8).copy(), w

    def codebook(cb, what):
        if len(cb) != 256:
            raise ValueError(f"{what} codebook: 256 floats expected, got {len(cb)}")
        return np.asarray(cb, np.float32)

    sheets = {}
    for k in ("lo", "hi", "scales"):
        sheets[k], _ = rgba(files[k], n)
    cbs = [codebook(scales["codebook"], "scales")]
    sheets["quats"], _ = rgba(files["quats"], n)
    tags = sheets["quats"].reshape(-1, 4)[:n, 3]
    bad = np.flatnonzero(tags < 252)
    if bad.size:
        raise ValueError(f"quats: invalid rotation tag {int(tags[bad[0]])}")
    sh_per_ch, cw = 1, 0
    if shn is not None:
        bands = shn["bands"]
        if bands not in (1, 2, 3):
            raise ValueError(f"shN.bands {bands} outside 1..=3")
        sh_per_ch = (bands + 1) ** 2
    sheets["sh0"], _ = rgba(files["sh0"], n)
    cbs.append(codebook(sh0["codebook"], "sh0"))
    if shn is not None:
        sheets["centroids"], cw = rgba(files["centroids"], 1)
        sheets["labels"], _ = rgba(files["labels"], n)
        cbs.append(codebook(shn["codebook"], "shN"))
        cpc = sh_per_ch - 1
        px = sheets["centroids"].size // 4
        count = px // cpc
        lb = sheets["labels"].reshape(-1, 4)[:n].astype(np.int64)
        label = lb[:, 0] | lb[:, 1] << 8
        per_row = cw // cpc
        if per_row:
            last = (label // per_row) * cw + (label % per_row) * cpc + cpc - 1
        else:
            last = np.full(n, px, np.int64)  # not one palette fits in a row
        bad = np.flatnonzero((label >= count) | (last >= px))
        if bad.size:
            i = bad[0]
            if label[i] >= count:
                raise ValueError(f"shN label {int(label[i])} >= palette size {count}")
            raise ValueError(f"shN label {int(label[i])}: its palette runs off the "
                             f"{cw}x{px // cw} centroid sheet")
    else:
        cbs.append(np.zeros(256, np.float32))
    info = dict(n=n, mins=[float(v) for v in means["mins"]],
                maxs=[float(v) for v in means["maxs"]], sh_per_ch=sh_per_ch)
    return info, sheets, cw, np.concatenate(cbs)
