"""Scene fixtures of the reference's test suites (src/render.rs:473-502,
src/seg/tests.rs:9-110, src/seg/views.rs:98-129), as numpy arrays."""
import numpy as np

from splatfield_amd.scene import ATTR_PLANES


def sample_opaque_attributes(n, position):
    """render.rs:478-494: identity rotation, log-scale -2, opacity logit 8."""
    a = np.zeros((ATTR_PLANES, n), np.float32)
    for i in range(n):
        a[0:3, i] = np.asarray(position(i), np.float32)
    a[3] = 1.0
    a[7:10] = -2.0
    a[10] = 8.0
    return a.reshape(-1)


def row_attributes(n, z, spacing):
    """seg/tests.rs:11-15."""
    f = np.float32
    return sample_opaque_attributes(
        n, lambda i: (f((f(i) - f(n - 1) * f(0.5)) * f(spacing)), 0.0, z))


def overdraw_attributes():
    """seg/tests.rs:78-89: 40 staggered splats crowding the centre tile."""
    f = np.float32

    def pos(i):
        x = f(i % 4) - f(1.5)
        y = f(i % 3) - f(1.0)
        return (x * f(0.04), y * f(0.04), f(1.0) + f(i) * f(0.5))
    return sample_opaque_attributes(40, pos)


def jitter(i):
    """views.rs:98-107."""
    f = np.float32
    return np.array([f((i * 73) % 17) / f(17.0) - f(0.5), f((i * 151) % 13) / f(13.0) - f(0.5),
                     f((i * 201) % 11) / f(11.0) - f(0.5)], np.float32)


# seg/tests.rs:186-203 — epsilon words of the overdraw scene at 70x70,
# produced by the reference's kernel on its CI device.
GOLDEN_EPS = [411225574, 64983936, 17882252, 5200864, 773093, 333408, 108187, 61572, 12526,
              2462, 1466, 1461, 10] + [0] * 27


def make_sog(n, bands, seed, width=64, bad_tag_at=None, bad_label_at=None):
    """A SOG archive (meta.json + lossless WebP sheets, the layout
    src/sog.rs:8-42 reads) of n random splats; sheets are width texels wide
    with rows padded past n.  Returns (zip bytes, the RGBA sheets)."""
    import io
    import json
    import zipfile
    from PIL import Image
    rng = np.random.default_rng(seed)
    rows = (n + width - 1) // width
    px = width * rows

    def sheet():
        return rng.integers(0, 256, (px, 4), dtype=np.uint8)
    sheets = {k: sheet() for k in ("lo", "hi", "scales", "quats", "sh0")}
    sheets["quats"][:, 3] = rng.integers(252, 256, px)
    if bad_tag_at is not None:
        sheets["quats"][bad_tag_at, 3] = 17
    cpc = (bands + 1) ** 2 - 1
    meta = {"count": n,
            "means": {"mins": [-1.5, -0.5, 0.25], "maxs": [1.25, 2.0, 3.0],
                      "files": ["means_l.webp", "means_u.webp"]},
            "scales": {"codebook": rng.normal(-4, 1, 256).tolist(), "files": ["scales.webp"]},
            "quats": {"files": ["quats.webp"]},
            "sh0": {"codebook": rng.normal(0, 0.5, 256).tolist(), "files": ["sh0.webp"]}}
    if bands:
        cw = 16 * cpc + 3  # a ragged row: the last 3 texels hold no palette
        crow = 7
        palettes = (cw // cpc) * crow
        sheets["centroids"] = rng.integers(0, 256, (cw * crow, 4), dtype=np.uint8)
        lab = rng.integers(0, palettes, px)
        if bad_label_at is not None:
            lab[bad_label_at] = palettes + 5
        sheets["labels"] = np.stack([lab & 255, lab >> 8, np.zeros(px), np.full(px, 255)],
                                    1).astype(np.uint8)
        meta["shN"] = {"bands": bands, "codebook": rng.normal(0, 0.1, 256).tolist(),
                       "files": ["shN_centroids.webp", "shN_labels.webp"]}
    names = {"lo": "means_l.webp", "hi": "means_u.webp", "scales": "scales.webp",
             "quats": "quats.webp", "sh0": "sh0.webp", "centroids": "shN_centroids.webp",
             "labels": "shN_labels.webp"}
    buf = io.BytesIO()
    with zipfile.ZipFile(buf, "w") as z:
        z.writestr("meta.json", json.dumps(meta))
        for k, v in sheets.items():
            w = cw if k == "centroids" else width
            img = Image.frombytes("RGBA", (w, v.shape[0] // w), v.tobytes())
            out = io.BytesIO()
            img.save(out, "WEBP", lossless=True, exact=True)
            z.writestr(names[k], out.getvalue())
    return buf.getvalue(), {k: v.reshape(-1) for k, v in sheets.items()}


PER_CLUSTER = 40  # views.rs:109-111


def cluster_scene():
    """views.rs:113-131: three well-separated spherical clusters of 40 opaque
    splats each; cluster 0 is the segmentation target.  Returns (attributes,
    [(centre, radius)])."""
    f = np.float32
    centers = [(np.array([1.2, 0.0, 0.0], f), f(0.30)),
               (np.array([-1.6, 0.6, 0.4], f), f(0.45)),
               (np.array([0.1, -0.4, 1.9], f), f(0.40))]

    def pos(i):
        c, r = centers[i // PER_CLUSTER]
        return c + (f(1.4) * r) * jitter(i)
    return sample_opaque_attributes(len(centers) * PER_CLUSTER, pos), centers


def _quat_matrix(q):
    x, y, z, w = (float(v) for v in q)
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def sphere_oracle(spheres):
    """views.rs:133-170: a perfect 2D oracle whose object is a set of
    spheres; a pixel is foreground iff its camera ray (the renderer's pixel
    convention, screen = focal * cam.xy / cam.z + size / 2) hits a sphere."""
    def oracle(rgb, size, camera):
        w, h = size
        mask = np.zeros((h, w), np.uint8)
        if not spheres:
            return mask.reshape(-1)
        fx, fy = camera.focal(w, h)
        xs = (np.arange(w) + 0.5 - w * 0.5) / float(fx)
        ys = (np.arange(h) + 0.5 - h * 0.5) / float(fy)
        X, Y = np.meshgrid(xs, ys)
        d = np.stack([X, Y, np.ones_like(X)], -1)
        d /= np.linalg.norm(d, axis=-1, keepdims=True)
        d = d @ _quat_matrix(camera.rotation).T
        pos = np.asarray(camera.position, np.float64)
        hit = np.zeros((h, w), bool)
        for c, r in spheres:
            oc = np.asarray(c, np.float64) - pos
            t = d @ oc
            l2 = float(oc @ oc)
            d2 = np.where(t < 0, l2, l2 - t * t)
            hit |= d2 <= float(r) * float(r)
        return hit.astype(np.uint8).reshape(-1)
    return oracle


def assert_image_close(rgba, final_t, bitmap, ref, n_contrib=None):
    """The RGB render contract against the oracle's exact evaluation: every
    f32 channel within 1/255 (the device takes exp(-power) from the hardware
    exp2), the mean error far below that (no systematic drift), the RGBA8
    bytes within one step, and the per-pixel last contributor equal on all
    but a sliver of pixels (a decision may flip only where alpha is within
    rounding of 1/255)."""
    d = np.abs(np.asarray(rgba, np.float64) - ref["rgba"])
    assert d.max() <= 1.0 / 255.0, d.max()
    assert d.mean() <= 1e-5, d.mean()
    dt = np.abs(np.asarray(final_t, np.float64) - ref["final_t"])
    assert dt.max() <= 1.0 / 255.0, dt.max()
    gb = np.asarray(bitmap, np.uint32).view(np.uint8).astype(np.int16)
    rb = np.asarray(ref["bitmap"], np.uint32).view(np.uint8).astype(np.int16)
    assert np.abs(gb - rb).max() <= 1
    if n_contrib is not None:
        assert np.mean(np.asarray(n_contrib) != ref["n_contrib"]) <= 1e-3
