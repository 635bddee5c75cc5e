"""The B3-Seg evidence round as specified for 1M splats (Beta(1,1) prior, a
footprint mask, a conjugate update, then 20 Fibonacci candidate views at
512^2 scored by analytic f64 EIG), on a 131,072-splat SH3 scene so that the
oracle checks every view: evidence words bit-exact, the update bit-exact,
each view's EIG within 1e-3 nats of the oracle's f64 sum over its own
words, and the batched round, the plain loop and the oracle agreeing on the
best view (src/seg/active.rs:110-183, src/seg/beta.rs:203-286)."""
import numpy as np
import pytest

from oracle import oracle as O
from splatfield_amd import render as R
from splatfield_amd import scene as S
from splatfield_amd import seg as G
from splatfield_amd.runtime import GpuTensor

pytestmark = pytest.mark.gpu

N, RES, VIEWS, THREADS = 131072, 512, 20, 16


def oview(v):
    return O.View.from_buffer_copy(bytes(v))


@pytest.fixture(scope="module")
def c3(ctx):
    attrs, sh, k = S.synthetic_scene(N, 3, seed=0)
    cam0 = R.Camera.look_at((0.0, 0.0, 3.0), (0.0, 0.0, 0.0))
    mask_bytes, subset = G.footprint_mask(ctx, attrs, sh, N, k, cam0, RES, RES, seed=0)
    assert 0 < subset.sum() < N and 0 < mask_bytes.sum() < RES * RES
    return dict(attrs=attrs, sh=sh, k=k, cam0=cam0, mask=mask_bytes)


def test_c3_observation_and_update(ctx, c3):
    """Round 0 from the bootstrap camera: the footprint mask's fg/bg words
    over the render's own isects equal the oracle's, and the Beta(1,1)
    update equals the oracle's a + fg/scale, b + bg/scale bit for bit."""
    attrs, sh, k = c3["attrs"], c3["sh"], c3["k"]
    splats = R.Splats(ctx, attrs, sh)
    scratch = R.RenderScratch(ctx, N, RES, RES)
    isects = R.prepare_isects(splats, scratch, c3["cam0"], RES, RES)
    R.finalize_rgb(scratch, isects)
    acc = G.Accumulators(ctx, N)
    mask = G.Mask(ctx, RES, RES, c3["mask"])
    G.accumulate_prepared(scratch, isects, acc, mask)
    scale = acc.scale
    assert scale == 2.0 ** 14
    ref = O.prepare(oview(c3["cam0"].view(RES, RES)), attrs, sh, N, k, RES, RES,
                    nthreads=THREADS)
    ev = O.rasterize(O.MODE_EVIDENCE, RES, RES, ref["ranges"], ref["gaussian_ids"],
                     ref["projected"], N, scale=scale, mask=mask.words,
                     words_per_row=mask.words_per_row, nthreads=THREADS)
    fg, bg = acc.fg.read(), acc.bg.read()
    np.testing.assert_array_equal(fg, ev["fg"], "fg words")
    np.testing.assert_array_equal(bg, ev["bg"], "bg words")
    assert fg.any() and bg.any()
    state = G.BetaState.new_uniform(ctx, N)
    state.update(acc)
    a, b = state.posteriors()
    ra, rb = O.beta_update(np.ones(N, np.float32), np.ones(N, np.float32), ev["fg"], ev["bg"],
                           scale)
    np.testing.assert_array_equal(a.view(np.uint32), ra.view(np.uint32), "a")
    np.testing.assert_array_equal(b.view(np.uint32), rb.view(np.uint32), "b")


def test_c3_round_against_oracle(ctx, c3):
    """20 candidate views after the round-0 update: every view's epsilon
    words of the batched round (hipGraph replay) equal the oracle's Total
    render, each view's f64 EIG is within 1e-3 nats of the oracle's sum over
    those words, the plain per-view loop scores the same, and the chosen
    view is the oracle's best (up to an EIG tie within the tolerance)."""
    attrs, sh, k = c3["attrs"], c3["sh"], c3["k"]
    splats = R.Splats(ctx, attrs, sh)
    state = G.observe_bootstrap(splats, c3["cam0"], RES, c3["mask"])
    a, b = state.posteriors()
    cams = G.sphere_cameras(VIEWS, radius=3.0)
    sizes = [(RES, RES)] * VIEWS
    rnd = G.EigRound(splats, cams, sizes, tile_w=16, digit_bits=8)
    eig_b, best_b = rnd.run(state, use_graph=True)
    eig_b, best_b = rnd.run(state, use_graph=True)  # a replay
    scratch = R.RenderScratch(ctx, N, RES, RES)
    resp = GpuTensor.empty(ctx, N)
    eig_p, best_p = G.eig_round_plain(splats, scratch, cams, sizes, state, resp)
    ref_eig = np.zeros(VIEWS)
    for v, cam in enumerate(cams):
        words, scale = rnd.resp(v)
        assert scale == 2.0 ** 14
        ref = O.prepare(oview(cam.view(RES, RES)), attrs, sh, N, k, RES, RES, nthreads=THREADS)
        tot = O.rasterize(O.MODE_TOTAL, RES, RES, ref["ranges"], ref["gaussian_ids"],
                          ref["projected"], N, scale=scale, nthreads=THREADS)
        np.testing.assert_array_equal(words, tot["resp"], f"view {v} epsilon words")
        ref_eig[v] = O.eig(a, b, tot["resp"], scale)
        assert abs(eig_b[v] - ref_eig[v]) <= 1e-3, (v, eig_b[v], ref_eig[v])
        assert abs(eig_p[v] - ref_eig[v]) <= 1e-3, (v, eig_p[v], ref_eig[v])
    assert best_b == best_p
    assert ref_eig[best_b] >= ref_eig.max() - 2e-3
    assert ref_eig.min() > 0
