"""Batched EIG round: per-view u32 evidence bit-exact against plain
single-view responsibility renders, EIG within 1e-3 nats of the plain f64
EIG and of the oracle, same selected view; hipGraph replays equal the eager
sequence; every tile width x digit width of the autotuner's search space."""
import numpy as np
import pytest

from oracle import oracle as O
from splatfield_amd import render as R
from splatfield_amd import scene as S
from splatfield_amd import seg as G
from splatfield_amd.runtime import GpuTensor

pytestmark = pytest.mark.gpu


def make_state(ctx, n, seed=4):
    rng = np.random.default_rng(seed)
    a = (1 + rng.random(n) * 6).astype(np.float32)
    b = (1 + rng.random(n) * 6).astype(np.float32)
    return G.BetaState(ctx, a, b), a, b


def plain_views(ctx, splats, cams, sizes, tile_w, state):
    out = []
    scratches = {}
    for cam, (w, h) in zip(cams, sizes):
        sc = scratches.get((w, h))
        if sc is None:
            sc = scratches[(w, h)] = R.RenderScratch(ctx, splats.n, w, h, tile_w=tile_w)
        acc = G.Accumulators(ctx, splats.n)
        G.render_responsibility(splats, sc, acc, cam, w, h)
        out.append((acc.bits.read(), acc.scale, state.eig(acc)))
    return out


@pytest.mark.parametrize("tile_w,digit_bits", [(16, 8), (8, 4), (32, 8), (16, 4), (8, 8), (32, 4)])
def test_batch_matches_plain(ctx, tile_w, digit_bits):
    n = 32768
    attrs, sh, k = S.synthetic_scene(n, 3, seed=2)
    splats = R.Splats(ctx, attrs, sh)
    cams = G.sphere_cameras(12)
    sizes = [(64, 64), (128, 128), (96, 80)] * 4
    state, a, b = make_state(ctx, n)
    rnd = G.EigRound(splats, cams, sizes, tile_w=tile_w, digit_bits=digit_bits)
    eig_e, best_e = rnd.run(state, use_graph=False)
    plain = plain_views(ctx, splats, cams, sizes, tile_w, state)
    for v, (bits, scale, e) in enumerate(plain):
        got, gscale = rnd.resp(v)
        assert gscale == scale
        np.testing.assert_array_equal(got, bits, f"view {v} evidence")
        assert abs(eig_e[v] - e) <= 1e-6 * max(1.0, abs(e))
        assert abs(eig_e[v] - O.eig(a, b, bits, scale)) <= 1e-3
    assert best_e == int(np.argmax([p[2] for p in plain]))
    # graph replays equal the eager sequence, round after round
    for _ in range(3):
        eig_g, best_g = rnd.run(state, use_graph=True)
        np.testing.assert_allclose(eig_g, eig_e, rtol=1e-12, atol=1e-12)
        assert best_g == best_e
        for v in (0, 5, 11):
            np.testing.assert_array_equal(rnd.resp(v)[0], plain[v][0])


def test_batch_c3_small_against_plain_loop(ctx):
    """The C3 shape at reduced size (65,536 splats, 20 views at 256^2):
    the C plain loop (select_view) and the batched round agree."""
    n = 65536
    attrs, sh, k = S.synthetic_scene(n, 3, seed=0)
    splats = R.Splats(ctx, attrs, sh)
    cams = G.sphere_cameras(20)
    sizes = [(256, 256)] * 20
    state, a, b = make_state(ctx, n, seed=9)
    scratch = R.RenderScratch(ctx, n, 256, 256)
    resp = GpuTensor.empty(ctx, n)
    eig_p, best_p = G.eig_round_plain(splats, scratch, cams, sizes, state, resp)
    rnd = G.EigRound(splats, cams, sizes)
    eig_b, best_b = rnd.run(state, use_graph=True)
    np.testing.assert_allclose(eig_b, eig_p, rtol=1e-9)
    assert best_b == best_p
    # the last plain view's words equal the batch's
    np.testing.assert_array_equal(rnd.resp(19)[0], resp.read())
    cnt = rnd.counters()
    assert cnt[-1] == cnt[0:-1:2].sum()


def test_batch_c4_full_mixed_resolution(ctx):
    """The C4 shape at full size: 2,097,152 splats (SH3), 64 Fibonacci-sphere
    cameras cycling 256^2 / 512^2 / 1024^2, one graph-captured batched round:
    every view's u32 evidence equals a plain render's bit for bit, the EIG
    values agree with the plain f64 EIG, and the same view is selected."""
    n = 1 << 21
    attrs, sh, k = S.synthetic_scene(n, 3, seed=0)
    splats = R.Splats(ctx, attrs, sh)
    cams = G.sphere_cameras(64, radius=3.0)
    res = (256, 512, 1024)
    sizes = [(res[i % 3], res[i % 3]) for i in range(64)]
    state, a, b = make_state(ctx, n, seed=11)
    rnd = G.EigRound(splats, cams, sizes)
    eig_b, best_b = rnd.run(state, use_graph=True)
    scratches = {}
    acc = G.Accumulators(ctx, n)
    eig_p = []
    for v, (cam, (w, h)) in enumerate(zip(cams, sizes)):
        sc = scratches.get(w)
        if sc is None:
            sc = scratches[w] = R.RenderScratch(ctx, n, w, h)
            sc.reserve_isects(1 << 25)  # no first-frame truncation
        G.render_responsibility(splats, sc, acc, cam, w, h)
        got, gscale = rnd.resp(v)
        assert gscale == acc.scale
        np.testing.assert_array_equal(got, acc.bits.read(), f"view {v}")
        eig_p.append(state.eig(acc))
    np.testing.assert_allclose(eig_b, eig_p, rtol=1e-9)
    assert best_b == int(np.argmax(eig_p))
    del rnd


@pytest.mark.parametrize("pair_raster", [True, False])
@pytest.mark.parametrize("tile_w", [16, 8, 32])
def test_batch_binned_matches_sorted(ctx, tile_w, pair_raster):
    """Batched binned emission (per-view supertile grids, ranges indexed
    local tile * views + view), rasterized straight from the supertile pairs
    or from placed per-tile lists, against the batched emission + stable tile
    argsort: every view's evidence words equal, eager and graph-replayed."""
    n = 32768
    attrs, sh, k = S.synthetic_scene(n, 0, seed=6)
    attrs = attrs.reshape(S.ATTR_PLANES, n)
    attrs[7:10] += np.float32(1.0)
    attrs = attrs.reshape(-1)
    splats = R.Splats(ctx, attrs, sh)
    cams = G.sphere_cameras(10)
    sizes = [(64, 64), (200, 120), (256, 256), (96, 80), (136, 136)] * 2
    state, a, b = make_state(ctx, n, seed=3)
    ref = G.EigRound(splats, cams, sizes, tile_w=tile_w)
    ref.set_binned(False)
    eig_s, best_s = ref.run(state, use_graph=False)
    words = [ref.resp(v)[0] for v in range(len(cams))]
    rnd = G.EigRound(splats, cams, sizes, tile_w=tile_w)
    rnd.set_pair_raster(pair_raster)
    for graph in (False, True, True):
        eig_b, best_b = rnd.run(state, use_graph=graph)
        assert best_b == best_s
        np.testing.assert_array_equal(eig_b, eig_s)
        for v in range(len(cams)):
            np.testing.assert_array_equal(rnd.resp(v)[0], words[v], f"view {v}")
    c_b, c_s = rnd.counters(), ref.counters()
    np.testing.assert_array_equal(c_b, c_s)


def test_batch_large_views_pair_supertiles(ctx):
    """A 2040 x 2040 view at 8 px tiles (255 x 255 tiles) holds more than
    256 of the pair emission's 64-pixel supertiles, so the round takes 128-
    pixel ones for it and stays on the direct pair emission; a wide view
    tests the row limit.  Evidence words and EIG equal the batched emission +
    stable tile argsort's, eager and graph-replayed."""
    n = 32768
    attrs, sh, k = S.synthetic_scene(n, 0, seed=9)
    attrs = attrs.reshape(S.ATTR_PLANES, n)
    attrs[7:10] += np.float32(1.0)
    attrs = attrs.reshape(-1)
    splats = R.Splats(ctx, attrs, sh)
    cams = G.sphere_cameras(3)
    sizes = [(2040, 2040), (2040, 64), (64, 64)]
    state, a, b = make_state(ctx, n, seed=5)
    ref = G.EigRound(splats, cams, sizes, tile_w=8)
    ref.set_binned(False)
    eig_s, best_s = ref.run(state, use_graph=True)
    words = [ref.resp(v)[0] for v in range(len(cams))]
    rnd = G.EigRound(splats, cams, sizes, tile_w=8)
    for graph in (True, True):
        eig_b, best_b = rnd.run(state, use_graph=graph)
        assert best_b == best_s
        np.testing.assert_array_equal(eig_b, eig_s)
        for v in range(len(cams)):
            np.testing.assert_array_equal(rnd.resp(v)[0], words[v], f"view {v}")


@pytest.mark.parametrize("n", [1 << 16, 1 << 20])
def test_batch_truncated_round(ctx, n):
    """A round whose isects exceed its capacity (an eager run before any
    capacity fit): the raster straight from the pairs drops exactly the
    isects whose emission slot reaches the capacity, as the placed lists and
    the emission + tile argsort do (raster.rs:12-48).  At 2^20 splats the
    (supertile, splat) pairs exceed the capacity too, and only those at
    rank-order slots below it are kept."""
    attrs, sh, k = S.synthetic_scene(n, 0, seed=8)
    attrs = attrs.reshape(S.ATTR_PLANES, n)
    attrs[7:10] += np.float32(2.0)  # large splats: ~10^8 isects at 8 px tiles
    attrs = attrs.reshape(-1)
    splats = R.Splats(ctx, attrs, sh)
    cams = G.sphere_cameras(2)
    sizes = [(1024, 1024), (512, 512)]
    state, a, b = make_state(ctx, n, seed=4)
    words = {}
    for mode in ("sorted", "placed", "pairs"):
        rnd = G.EigRound(splats, cams, sizes, tile_w=8)
        rnd.set_binned(mode != "sorted")
        rnd.set_pair_raster(mode == "pairs")
        rnd.run(state, use_graph=False)
        c = rnd.counters()
        assert c[-1] > rnd.capacity(), "the round must truncate"
        if n == 1 << 20 and mode == "pairs":
            # (supertile, splat) pairs of the two views, from the oracle's
            # tile bboxes: more than the capacity holds.
            pairs = 0
            for cam, (w, h) in zip(cams, sizes):
                view = O.View.from_buffer_copy(bytes(cam.view(w, h, 8)))
                bb = O.project(view, attrs, sh, n, k, 8, nthreads=8)["bbox"].reshape(-1, 2)
                x0, y0 = bb[:, 0] & 0xffff, bb[:, 0] >> 16
                x1, y1 = bb[:, 1] & 0xffff, bb[:, 1] >> 16
                live = (x1 > x0) & (y1 > y0)
                sx = (x1[live] - 1) // 8 - x0[live] // 8 + 1
                sy = (y1[live] - 1) // 8 - y0[live] // 8 + 1
                pairs += int((sx.astype(np.int64) * sy).sum())
            assert pairs > rnd.capacity(), "the pairs must exceed the capacity"
        words[mode] = [rnd.resp(v)[0] for v in range(len(cams))]
        del rnd
    for mode in ("placed", "pairs"):
        for v in range(len(cams)):
            np.testing.assert_array_equal(words[mode][v], words["sorted"][v], f"{mode} view {v}")
