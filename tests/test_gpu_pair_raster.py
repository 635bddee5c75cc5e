"""Pair raster (sfk_render_scratch_set_pair_raster): the binned emission stops
before placing the isects, and each rasterizer takes a tile's isects straight
from its supertile's depth-ordered pairs, writing back the ones it consumes.
Every output equals the placed lists' -- RGBA and final T bit for bit,
n_contrib (absolute isect slots), the RGBA8 bitmap, total and evidence words,
the consumed isect ids -- and the backward pass reads the written-back ids:
its gradients equal the placed path's up to the order of the float atomics
that gather them across tiles."""
import math

import numpy as np
import pytest

from splatfield_amd import render as R
from splatfield_amd import scene as S
from splatfield_amd import seg as G
from splatfield_amd.runtime import GpuTensor

pytestmark = pytest.mark.gpu


def _frame(ctx, splats, n, k, view, w, h, tile_w, pair, target=None, mask=None):
    scratch = R.RenderScratch(ctx, n, w, h, tile_w=tile_w)
    scratch.set_pair_raster(pair)
    if pair:
        scratch.set_tile_ids(False)
    isects = R.prepare_isects(splats, scratch, view=view, w=w, h=h)
    R.finalize_rgb(scratch, isects)
    out = {k_: scratch.read(k_) for k_ in ("rgba", "final_t", "n_contrib", "bitmap")}
    out["starts"] = isects.tile_starts()
    out["gids"] = isects.gaussian_ids(isects.counts()[2])
    acc = G.Accumulators(ctx, n)
    G.accumulate_prepared(scratch, isects, acc)
    out["total"] = acc.bits.read()
    if mask is not None:
        G.accumulate_prepared(scratch, isects, acc, G.Mask(ctx, w, h, mask))
        out["fg"], out["bg"] = acc.fg.read(), acc.bg.read()
    if target is not None:
        # the forward again (the responsibility renders left the RGB state)
        R.finalize_rgb(scratch, isects)
        tgt = GpuTensor.from_numpy(ctx, target)
        ga = GpuTensor.empty(ctx, n * 11, np.float32)
        gs = GpuTensor.empty(ctx, n * 3 * k, np.float32)
        loss = GpuTensor.empty(ctx, 1, np.float64)
        R.render_backward_l2(splats, scratch, isects, scratch.device("rgba"), tgt.ptr, ga, gs,
                             loss)
        out["ga"], out["gs"], out["loss"] = ga.read(), gs.read(), float(loss.read()[0])
    out["counts"] = isects.counts()
    return out


def _consumed(out, tile_w, w, h):
    """Per tile, the isect ids up to the last contributor of any pixel."""
    tx = (w + tile_w - 1) // tile_w
    nc = out["n_contrib"].reshape(h, w)
    ids = []
    for t in range(len(out["starts"]) - 1):
        x, y = t % tx, t // tx
        last = int(nc[y * tile_w:(y + 1) * tile_w, x * tile_w:(x + 1) * tile_w].max())
        ids.append(out["gids"][out["starts"][t]:last])
    return ids


def _compare(a, b, tile_w, w, h):
    for key in ("rgba", "final_t"):
        np.testing.assert_array_equal(a[key].view(np.uint32), b[key].view(np.uint32), key)
    for key in ("n_contrib", "bitmap", "starts", "total"):
        np.testing.assert_array_equal(a[key], b[key], key)
    for key in ("fg", "bg"):
        if key in a:
            np.testing.assert_array_equal(a[key], b[key], key)
    assert tuple(a["counts"]) == tuple(b["counts"])
    for t, (x, y) in enumerate(zip(_consumed(a, tile_w, w, h), _consumed(b, tile_w, w, h))):
        np.testing.assert_array_equal(x, y, f"consumed ids of tile {t}")


def _camera(w, h):
    fy = np.float32(math.pi / 3)
    fx = np.float32(S.fov_for_aspect(w, h, fy))
    return R.Camera.look_at((0.0, 0.0, 3.0), (0.0, 0.0, 0.0), fov=(fx, fy))


@pytest.mark.parametrize("n,sh_deg,w,h,tile_w,boost", [
    (65536, 0, 256, 256, 16, 0.0),
    (16384, 3, 320, 180, 16, 0.0),
    (16384, 3, 320, 180, 8, 0.0),
    (16384, 3, 320, 180, 32, 0.0),
    (65536, 0, 1920, 1080, 16, 1.6),   # truncated first frame
])
def test_pair_raster_matches_placed(ctx, n, sh_deg, w, h, tile_w, boost):
    attrs, sh, k = S.synthetic_scene(n, sh_deg, seed=31)
    if boost:
        attrs = attrs.reshape(S.ATTR_PLANES, n)
        attrs[7:10] += np.float32(boost)
        attrs = attrs.reshape(-1)
    splats = R.Splats(ctx, attrs, sh)
    view = _camera(w, h).view(w, h, tile_w)
    rng = np.random.default_rng(5)
    mask = (rng.random(w * h) < 0.3).astype(np.uint8)
    placed = _frame(ctx, splats, n, k, view, w, h, tile_w, False, mask=mask)
    pair = _frame(ctx, splats, n, k, view, w, h, tile_w, True, mask=mask)
    if boost:
        assert placed["counts"][0] > placed["counts"][2], "the first frame must truncate"
    _compare(pair, placed, tile_w, w, h)


@pytest.mark.parametrize("n,w,h", [(4096, 64, 64), (262144, 960, 540)])
def test_pair_raster_backward(ctx, n, w, h):
    attrs, sh, k = S.synthetic_scene(n, 3, seed=21)
    rng = np.random.default_rng(22)
    pa = (attrs + rng.normal(0, 0.01, attrs.size)).astype(np.float32)
    ps = (sh + rng.normal(0, 0.01, sh.size)).astype(np.float32)
    view = _camera(w, h).view(w, h)
    sp_t = R.Splats(ctx, pa, ps)
    sc_t = R.RenderScratch(ctx, n, w, h)
    R.finalize_rgb(sc_t, R.prepare_isects(sp_t, sc_t, view=view, w=w, h=h))
    target = sc_t.read("rgba")
    splats = R.Splats(ctx, attrs, sh)
    placed = _frame(ctx, splats, n, k, view, w, h, 16, False, target=target)
    pair = _frame(ctx, splats, n, k, view, w, h, 16, True, target=target)
    _compare(pair, placed, 16, w, h)
    assert pair["loss"] == placed["loss"]
    # Float atomics gather each splat's gradient across tiles in no fixed
    # order: elements that nearly cancel differ by a few ulps of the terms.
    for key in ("ga", "gs"):
        g, r = pair[key], placed[key]
        np.testing.assert_allclose(g, r, rtol=1e-4, atol=1e-5 * float(np.abs(r).max()), err_msg=key)
        assert np.linalg.norm(g - r) <= 1e-6 * np.linalg.norm(r), key
