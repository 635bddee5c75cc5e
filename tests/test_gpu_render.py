"""GPU parity of the render path (project -> depth sort -> scan -> map ->
tile sort -> ranges -> raster) against the oracle, on the same inputs.

Integer outputs (visibility, depth keys, tile counts, bboxes, counters, sort
order, offsets, tile ranges, u32 epsilon) must be bit-exact.  The RGB image
is held to the contract's 1/255 per channel (the device's RGB compositor
takes exp(-power) from the hardware exp2; the responsibility modes keep the
exact exp their u32 words are specified with)."""
import numpy as np
import pytest

from oracle import oracle as O
from splatfield_amd import render as R
from splatfield_amd import scene as S
from splatfield_amd import seg as G
from splatfield_amd.runtime import GpuTensor

from .fixtures import (GOLDEN_EPS, assert_image_close, overdraw_attributes,
                       sample_opaque_attributes)

pytestmark = pytest.mark.gpu


def oview(v):
    return O.View.from_buffer_copy(bytes(v))


def c1_camera():
    # C1: camera at (0,0,3) looking at the origin, 60 deg fov on both axes.
    fov = np.float32(np.pi / 3)
    return R.Camera.look_at((0.0, 0.0, 3.0), (0.0, 0.0, 0.0), fov=(fov, fov))


def run_both(ctx, attrs, sh, n, per_ch, cam, w, h, tile_w=16, reserve=None):
    splats = R.Splats(ctx, attrs, sh)
    scratch = R.RenderScratch(ctx, n, w, h, tile_w=tile_w)
    if reserve:
        scratch.reserve_isects(reserve)
    view = cam.view(w, h, tile_w)
    isects = R.prepare_isects(splats, scratch, view=view, w=w, h=h)
    ref = O.prepare(oview(view), attrs, sh, n, per_ch, w, h, tile_w)
    return splats, scratch, view, isects, ref


def check_prepare(scratch, isects, ref, n):
    keys = scratch.read("depth_keys")
    vis = keys != 0xFFFFFFFF
    np.testing.assert_array_equal(vis, ref["visible"].astype(bool), "visibility")
    np.testing.assert_array_equal(keys[vis], ref["depth_keys"][vis], "depth keys")
    proj = scratch.read("projected").reshape(n, R.PROJ_STRIDE)
    rp = ref["projected"].reshape(n, 9)
    got = proj[:, [0, 1, 2, 3, 4, 6, 7, 8, 5]]
    np.testing.assert_array_equal(got[vis].view(np.uint32), rp[vis].view(np.uint32),
                                  "projected rows")
    tc = scratch.read("tile_counts")
    np.testing.assert_array_equal(tc[vis], ref["tile_counts"][vis], "tile counts")
    bb = scratch.read("bbox").reshape(n, 2)
    np.testing.assert_array_equal(bb[vis], ref["bbox"].reshape(n, 2)[vis], "bbox")
    raw, nvis, ni = isects.counts()
    assert (raw, nvis) == tuple(int(x) for x in ref["counters"]), "counters"
    assert ni == ref["num_isects"]
    order = scratch.read("depth_order")
    np.testing.assert_array_equal(order[:nvis], ref["depth_order"], "depth order")
    offs = scratch.read("offsets")
    np.testing.assert_array_equal(offs[:nvis], ref["offsets"], "emission offsets")
    tid = isects.tile_ids(ni)
    gid = isects.gaussian_ids(ni)
    np.testing.assert_array_equal(tid, ref["tile_ids"], "sorted tile ids")
    np.testing.assert_array_equal(gid, ref["gaussian_ids"], "sorted splat ids")
    starts = isects.tile_starts()
    T = ref["tiles_x"] * ref["tiles_y"]
    want = np.array([O.lower_bound(ref["tile_ids"], t) for t in range(T + 1)], np.uint32) \
        if T < 4096 else np.concatenate([ref["ranges"][0::2], [ni]]).astype(np.uint32)
    np.testing.assert_array_equal(starts, want, "tile ranges")
    return ni


def check_rgb(scratch, isects, ref, n, w, h, tile_w):
    R.finalize_rgb(scratch, isects)
    ras = O.rasterize(O.MODE_RGB, w, h, ref["ranges"], ref["gaussian_ids"], ref["projected"], n,
                      tile_w)
    assert_image_close(scratch.read("rgba"), scratch.read("final_t"), scratch.read("bitmap"), ras,
                       scratch.read("n_contrib"))
    return ras


def check_eps(ctx, scratch, isects, ref, n, w, h, tile_w, mask_bytes=None):
    fb = G.fractional_bits(w, h)
    scale = float(2 ** fb)
    acc = G.Accumulators(ctx, n)
    G.accumulate_prepared(scratch, isects, acc)
    assert acc.scale == scale
    tot = O.rasterize(O.MODE_TOTAL, w, h, ref["ranges"], ref["gaussian_ids"], ref["projected"], n,
                      tile_w, scale=scale)
    np.testing.assert_array_equal(acc.bits.read(), tot["resp"], "total epsilon words")
    if mask_bytes is not None:
        m = G.Mask(ctx, w, h, mask_bytes)
        G.accumulate_prepared(scratch, isects, acc, m)
        ev = O.rasterize(O.MODE_EVIDENCE, w, h, ref["ranges"], ref["gaussian_ids"],
                         ref["projected"], n, tile_w, scale=scale, mask=m.words,
                         words_per_row=m.words_per_row)
        fg, bg = acc.fg.read(), acc.bg.read()
        np.testing.assert_array_equal(fg, ev["fg"], "fg words")
        np.testing.assert_array_equal(bg, ev["bg"], "bg words")
        np.testing.assert_array_equal(fg + bg, tot["resp"], "fg + bg == total")
    return acc


def test_project_kat(ctx):
    """project.rs:399-446: one splat, 64x64, covers all 16 tiles."""
    attrs = np.array([0, 0, 1, 1, 0, 0, 0, 0, 0, 0, 5], np.float32)
    sh = np.zeros(3, np.float32)
    splats, scratch, view, isects, ref = run_both(ctx, attrs, sh, 1, 1, R.Camera(), 64, 64)
    raw, nvis, _ = isects.counts()
    assert (raw, nvis) == (16, 1)
    assert scratch.read("tile_counts")[0] == 16
    np.testing.assert_array_equal(scratch.read("bbox"), [0, 4 | (4 << 16)])
    check_prepare(scratch, isects, ref, 1)


@pytest.mark.parametrize("tile_w", [16, 8, 32])
def test_c1_parity(ctx, tile_w):
    """C1: 65,536 splats, SH0, 256x256, seed 0: every stage bit-exact."""
    n = 65536
    attrs, sh, k = S.synthetic_scene(n, 0, seed=0)
    splats, scratch, view, isects, ref = run_both(ctx, attrs, sh, n, k, c1_camera(), 256, 256,
                                                  tile_w)
    ni = check_prepare(scratch, isects, ref, n)
    assert ni > 0
    check_rgb(scratch, isects, ref, n, 256, 256, tile_w)
    mask = np.zeros((256, 256), np.uint8)
    mask[:, :128] = 1
    check_eps(ctx, scratch, isects, ref, n, 256, 256, tile_w, mask.reshape(-1))
    # sorted u64 tile|depth keys are non-decreasing and match the oracle's
    keys = O.tile_keys_u64(isects.tile_ids(ni), isects.gaussian_ids(ni), scratch.read("depth_keys"))
    assert np.all(keys[1:] >= keys[:-1])
    np.testing.assert_array_equal(keys, O.tile_keys_u64(ref["tile_ids"], ref["gaussian_ids"],
                                                        ref["depth_keys"]))


def test_sh3_parity(ctx):
    """SH degree 3 colours through the GPU projection, bit-exact (C2 sizes
    scaled to 16,384 splats at 320x180)."""
    n = 16384
    attrs, sh, k = S.synthetic_scene(n, 3, seed=3)
    fy = np.float32(np.pi / 3)
    fx = np.float32(S.fov_for_aspect(320, 180, fy))
    cam = R.Camera.look_at((0.0, 0.0, 3.0), (0.0, 0.0, 0.0), fov=(fx, fy))
    splats, scratch, view, isects, ref = run_both(ctx, attrs, sh, n, k, cam, 320, 180)
    check_prepare(scratch, isects, ref, n)
    check_rgb(scratch, isects, ref, n, 320, 180, 16)


def thin_rotated_scene(n, seed, long_big=(-0.6, 0.2)):
    """Needle-like splats (one long axis, two thin ones; a tenth with the long
    axis drawn from `long_big`, which by default gives conic condition numbers
    past 1e4), random rotations, opacities from just above the 10/255 cull to
    nearly opaque: the shapes where a splat's reach box covers many 8x8
    blocks its alpha >= 1/255 ellipse misses."""
    rng = np.random.default_rng(seed)
    a = np.zeros((11, n), np.float32)
    a[0] = rng.uniform(-0.9, 0.9, n)
    a[1] = rng.uniform(-0.7, 0.7, n)
    a[2] = rng.uniform(-0.5, 0.5, n)
    q = rng.normal(size=(4, n))
    a[3:7] = q / np.linalg.norm(q, axis=0)
    a[7] = np.where(rng.random(n) < 0.1, rng.uniform(*long_big, n), rng.normal(-2.0, 0.3, n))
    thin = np.where(rng.random(n) < 0.2, -7.5, -5.5)
    a[8] = thin + rng.normal(0, 0.2, n)
    a[9] = thin + rng.normal(0, 0.2, n)
    a[10] = rng.uniform(-3.2, 4.0, n)
    sh = rng.normal(0, 0.3, (1, 3, n)).astype(np.float32)
    return a.reshape(-1), sh.reshape(-1)


@pytest.mark.parametrize("tile_w", [16, 8])
def test_thin_rotated_splats_parity(ctx, tile_w):
    """The per-wave splat lists drop a splat only where its least power over
    the block's pixel centres exceeds tau: on needle-shaped, rotated splats
    (the case that test exercises most) every stage stays bit-exact."""
    n, w, h = 3000, 256, 192
    attrs, sh = thin_rotated_scene(n, seed=7)
    splats, scratch, view, isects, ref = run_both(ctx, attrs, sh, n, 1, c1_camera(), w, h, tile_w)
    ni = check_prepare(scratch, isects, ref, n)
    assert ni > 0
    check_rgb(scratch, isects, ref, n, w, h, tile_w)
    mask = np.zeros((h, w), np.uint8)
    mask[:, : w // 2] = 1
    check_eps(ctx, scratch, isects, ref, n, w, h, tile_w, mask.reshape(-1))


def test_golden_center(ctx):
    """render.rs:671-710: stacked opaque splats, centre 127 +-1, alpha >= 254."""
    n = 5
    s2 = np.float32(np.sin(np.float32(np.pi / 4)))
    c2 = np.float32(np.cos(np.float32(np.pi / 4)))
    cases = [
        (sample_opaque_attributes(n, lambda i: (0, 0, 1 + i * 0.5)), R.Camera()),
        (sample_opaque_attributes(n, lambda i: (0, 0, 0)),
         R.Camera(position=(-2.0, 0.0, 0.0), rotation=(0.0, s2, 0.0, c2))),
    ]
    for attrs, cam in cases:
        splats = R.Splats(ctx, attrs, np.zeros(3 * n, np.float32))
        scratch = R.RenderScratch(ctx, n, 32, 32)
        bm = R.render_with(splats, scratch, cam, 32, 32)
        c = int(bm[16, 16])
        r, g, b, a = c & 255, (c >> 8) & 255, (c >> 16) & 255, c >> 24
        assert a >= 254
        assert int(bm[0, 0]) & 0xFF000000 == 0
        for v in (r, g, b):
            assert abs(v - 127) <= 1


def test_eps_golden_overdraw(ctx):
    """seg/tests.rs:186-203: the 40 GOLDEN epsilon words of the overdraw
    scene.  GOLDEN came from the reference's device exp; ours must match
    the oracle bit-exactly and GOLDEN within the exp ulp budget."""
    attrs = overdraw_attributes()
    n = 40
    sh = np.zeros(3 * n, np.float32)
    splats, scratch, view, isects, ref = run_both(ctx, attrs, sh, n, 1, R.Camera(), 70, 70)
    check_prepare(scratch, isects, ref, n)
    acc = check_eps(ctx, scratch, isects, ref, n, 70, 70, 16)
    bits = acc.bits.read().astype(np.int64)
    gold = np.array(GOLDEN_EPS, np.int64)
    assert np.all((bits == 0) == (gold == 0))
    assert np.all(np.abs(bits - gold) <= np.maximum(200, gold * 1e-6))


def test_ties_deterministic(ctx):
    """render.rs:784-812: 200 identical splats at one depth."""
    n = 200
    attrs = sample_opaque_attributes(n, lambda i: (0, 0, 1))
    splats = R.Splats(ctx, attrs, np.zeros(3 * n, np.float32))
    s1 = R.RenderScratch(ctx, n, 64, 64)
    a = R.render_with(splats, s1, R.Camera(), 64, 64)
    b = R.render_with(splats, s1, R.Camera(), 64, 64)
    s2 = R.RenderScratch(ctx, n, 64, 64)
    c = R.render_with(splats, s2, R.Camera(), 64, 64)
    np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(a, c)
    # ties rank by splat index
    order = s1.read("depth_order")
    np.testing.assert_array_equal(order, np.arange(n))


def _isect_arrays(scratch, isects):
    raw, nvis, ni = isects.counts()
    return dict(counts=(raw, nvis, ni), tile_ids=isects.tile_ids(ni),
                gids=isects.gaussian_ids(ni), starts=isects.tile_starts(),
                order=scratch.read("depth_order")[:nvis], offsets=scratch.read("offsets")[:nvis])


@pytest.mark.parametrize("w,h,tile_w,n,boost", [
    (256, 256, 16, 65536, 0.0),    # 16 x 16 tiles: 4 supertiles
    (320, 180, 16, 16384, 1.0),    # 20 x 12 tiles: partial supertiles
    (64, 64, 16, 2048, 1.5),       # one supertile (a 1-bit pair sort)
    (1920, 1080, 8, 65536, 0.5),   # 240 x 135 tiles: 510 supertiles, two pair passes
    (1024, 768, 32, 32768, 1.0),   # 32 x 24 tiles
])
def test_binned_matches_sorted(ctx, w, h, tile_w, n, boost):
    """The binned emission (supertile pairs, per-tile counts from a wave bit
    transpose, deterministic placement) against emission + stable tile
    argsort: every isect array word for word."""
    attrs, sh, k = S.synthetic_scene(n, 0, seed=11)
    attrs = attrs.reshape(S.ATTR_PLANES, n)
    attrs[7:10] += np.float32(boost)  # larger splats: more tiles and supertiles each
    attrs = attrs.reshape(-1)
    fy = np.float32(np.pi / 3)
    fx = np.float32(S.fov_for_aspect(w, h, fy))
    cam = R.Camera.look_at((0.0, 0.0, 3.0), (0.0, 0.0, 0.0), fov=(fx, fy))
    splats = R.Splats(ctx, attrs, sh)
    out = []
    for binned in (True, False):
        scratch = R.RenderScratch(ctx, n, w, h, tile_w=tile_w)
        scratch.set_binned(binned)
        scratch.reserve_isects(1 << 24)
        view = cam.view(w, h, tile_w)
        out.append(_isect_arrays(scratch, R.prepare_isects(splats, scratch, view=view, w=w, h=h)))
    a, b = out
    assert a["counts"] == b["counts"] and a["counts"][2] > 0
    for key in ("tile_ids", "gids", "starts", "order", "offsets"):
        np.testing.assert_array_equal(a[key], b[key], key)


def test_binned_without_tile_ids(ctx):
    """With tile ids switched off the binned emission writes the same splat
    ids and tile ranges, the RGB render is unchanged bit for bit, and the
    tile ids are reported absent; switched back on, they are written again."""
    n, w, h = 16384, 320, 180
    attrs, sh, k = S.synthetic_scene(n, 1, seed=12)
    cam = R.Camera.look_at((0.0, 0.0, 3.0), (0.0, 0.0, 0.0))
    view = cam.view(w, h)
    splats = R.Splats(ctx, attrs, sh)
    scratch = R.RenderScratch(ctx, n, w, h)
    scratch.reserve_isects(1 << 22)
    isects = R.prepare_isects(splats, scratch, view=view, w=w, h=h)
    full = _isect_arrays(scratch, isects)
    R.finalize_rgb(scratch, isects)
    rgba = scratch.read("rgba").copy()
    scratch.set_tile_ids(False)
    isects = R.prepare_isects(splats, scratch, view=view, w=w, h=h)
    ni = isects.counts()[2]
    assert ni == full["counts"][2] > 0
    np.testing.assert_array_equal(isects.gaussian_ids(ni), full["gids"])
    np.testing.assert_array_equal(isects.tile_starts(), full["starts"])
    with pytest.raises(ValueError):
        isects.tile_ids(ni)
    R.finalize_rgb(scratch, isects)
    np.testing.assert_array_equal(scratch.read("rgba"), rgba)
    scratch.set_tile_ids(True)
    isects = R.prepare_isects(splats, scratch, view=view, w=w, h=h)
    np.testing.assert_array_equal(isects.tile_ids(ni), full["tile_ids"])


def test_truncated_first_frame(ctx):
    """A fresh scratch holds min(tiles * n, 2^22) isects (render.rs:143-167);
    a first frame emitting more keeps exactly the isects whose emission slot
    is below that (raster.rs:12-48), in both emission paths, as the oracle
    does with the same max_isects."""
    n, w, h = 65536, 1920, 1080
    attrs, sh, k = S.synthetic_scene(n, 0, seed=5)
    attrs = attrs.reshape(S.ATTR_PLANES, n)
    attrs[7:10] += np.float32(1.6)
    attrs = attrs.reshape(-1)
    fy = np.float32(np.pi / 3)
    fx = np.float32(S.fov_for_aspect(w, h, fy))
    cam = R.Camera.look_at((0.0, 0.0, 3.0), (0.0, 0.0, 0.0), fov=(fx, fy))
    view = cam.view(w, h)
    cap = 1 << 22
    ref = O.prepare(oview(view), attrs, sh, n, k, w, h, 16, max_isects=cap, nthreads=8)
    assert ref["num_isects_raw"] > cap, "scene must overflow the first-frame capacity"
    splats = R.Splats(ctx, attrs, sh)
    for binned in (True, False):
        scratch = R.RenderScratch(ctx, n, w, h)
        scratch.set_binned(binned)
        isects = R.prepare_isects(splats, scratch, view=view, w=w, h=h)
        assert isects.counts()[2] == cap
        check_prepare(scratch, isects, ref, n)


@pytest.mark.parametrize("zspan", [(-1.0, 1.0), (-40.0, 2.0)])
def test_depth_sort_key_window(ctx, zspan):
    """The depth sort runs on 24 bits of keys rebased to the view's depth
    window when the splats' AABB confines the visible depths to 2^24 f32 bit
    patterns (camera 3 units from a [-1,1]^3 scene), and on all 32 bits
    otherwise (depths 1..43): both orders equal the oracle's, ties and
    culled splats included."""
    n = 16384
    attrs, sh, k = S.synthetic_scene(n, 0, seed=21)
    attrs = attrs.reshape(S.ATTR_PLANES, n)
    rng = np.random.default_rng(21)
    attrs[2] = rng.uniform(zspan[0], zspan[1], n).astype(np.float32)
    attrs[2, :64] = attrs[2, 64:128]   # exact depth ties
    attrs[0, -32:] = 500.0             # off-screen: culled
    attrs = attrs.reshape(-1)
    cam = R.Camera.look_at((0.0, 0.0, 3.0), (0.0, 0.0, 0.0))
    splats, scratch, view, isects, ref = run_both(ctx, attrs, sh, n, k, cam, 320, 240)
    check_prepare(scratch, isects, ref, n)
    check_rgb(scratch, isects, ref, n, 320, 240, 16)


def test_c2_full_frame_against_oracle(ctx):
    """The full-size forward frame (1,048,576 SH3 splats, 1920x1080, seed 0,
    the benchmark's camera): visibility, keys, rows, counts, bboxes,
    counters, depth order, per-rank offsets, sorted tile and splat ids and
    every tile range equal to the oracle's (~9.8M isects, 13 tile bits,
    24-bit rebased depth sort, binned emission); the image within 1/255."""
    n, w, h = 1 << 20, 1920, 1080
    attrs, sh, k = S.synthetic_scene(n, 3, seed=0)
    fy = np.float32(np.pi / 3)
    fx = np.float32(S.fov_for_aspect(w, h, fy))
    cam = R.Camera.look_at((0.0, 0.0, 3.0), (0.0, 0.0, 0.0), fov=(fx, fy))
    splats = R.Splats(ctx, attrs, sh)
    scratch = R.RenderScratch(ctx, n, w, h)
    scratch.reserve_isects(16 << 20)
    view = cam.view(w, h)
    isects = R.prepare_isects(splats, scratch, view=view, w=w, h=h)
    ref = O.prepare(oview(view), attrs, sh, n, k, w, h, 16, nthreads=16)
    assert ref["num_isects"] > 8_000_000
    ni = check_prepare(scratch, isects, ref, n)
    assert ni == ref["num_isects_raw"]
    R.finalize_rgb(scratch, isects)
    ras = O.rasterize(O.MODE_RGB, w, h, ref["ranges"], ref["gaussian_ids"], ref["projected"], n,
                      16, nthreads=16)
    assert_image_close(scratch.read("rgba"), scratch.read("final_t"), scratch.read("bitmap"), ras,
                       scratch.read("n_contrib"))
