"""GPU B3-Seg evidence path against the oracle, following the reference's
suites (src/seg/tests.rs, src/seg/beta.rs tests).  u32 evidence words are
bit-exact; EIG is within 1e-3 nats of the oracle's f64 sum."""
import numpy as np
import pytest

from oracle import oracle as O
from splatfield_amd import render as R
from splatfield_amd import scene as S
from splatfield_amd import seg as G
from splatfield_amd.runtime import GpuTensor

from .fixtures import overdraw_attributes, row_attributes, sample_opaque_attributes

pytestmark = pytest.mark.gpu


def opaque(ctx, attrs):
    n = attrs.size // 11
    return R.Splats(ctx, attrs, np.zeros(3 * n, np.float32)), n


def left_half_mask(ctx, w, h):
    b = np.zeros((h, w), np.uint8)
    b[:, : w // 2] = 1
    return G.Mask(ctx, w, h, b.reshape(-1))


def test_responsibility_conserves_pixel_weights(ctx):
    """seg/tests.rs:112-126."""
    sp, n = opaque(ctx, sample_opaque_attributes(5, lambda i: (0, 0, 1 + 0.5 * i)))
    sc = R.RenderScratch(ctx, n, 64, 64)
    bm = R.render_with(sp, sc, R.Camera(), 64, 64)
    alpha = (bm[:, :64] >> 24).astype(np.float64).sum() / 255
    acc = G.Accumulators(ctx, n)
    G.render_responsibility(sp, sc, acc, R.Camera(), 64, 64)
    assert abs(alpha - acc.read_f32().sum()) < 64 * 64 * 3 / 255


def test_responsibility_orders_depth_and_culls(ctx):
    """seg/tests.rs:128-150."""
    sp, n = opaque(ctx, sample_opaque_attributes(
        6, lambda i: (0, 0, -3 if i == 5 else 1 + 0.5 * i)))
    sc = R.RenderScratch(ctx, n, 64, 64)
    acc = G.Accumulators(ctx, n)
    G.render_responsibility(sp, sc, acc, R.Camera(), 64, 64)
    eps = acc.read_f32()
    assert all(eps[k] > eps[k + 1] for k in range(4))
    assert eps[5] == 0


def test_responsibility_bit_deterministic(ctx):
    """seg/tests.rs:152-179."""
    sp, n = opaque(ctx, overdraw_attributes())
    sc = R.RenderScratch(ctx, n, 70, 70)
    acc = G.Accumulators(ctx, n)
    ref = None
    for _ in range(8):
        G.render_responsibility(sp, sc, acc, R.Camera(), 70, 70)
        bits = acc.bits.read()
        if ref is None:
            assert bits.any()
            ref = bits
        np.testing.assert_array_equal(bits, ref)


def split_scene(ctx):
    sp, n = opaque(ctx, row_attributes(2, 2.0, 1.0))
    sc = R.RenderScratch(ctx, n, 64, 64)
    acc = G.Accumulators(ctx, n)
    G.render_responsibility(sp, sc, acc, R.Camera(), 64, 64)
    mask = left_half_mask(ctx, 64, 64)
    it = R.prepare_isects(sp, sc, R.Camera(), 64, 64)
    G.accumulate_prepared(sc, it, acc, mask)
    return sp, sc, acc


def test_evidence_splits_responsibility_exactly(ctx):
    """seg/tests.rs:205-227."""
    _, _, acc = split_scene(ctx)
    fg, bg, tot = acc.fg.read(), acc.bg.read(), acc.bits.read()
    np.testing.assert_array_equal(fg + bg, tot)
    assert bg[0] == 0 and fg[1] == 0
    f, b = acc.read_evidence_f32()
    e = acc.read_f32()
    assert f[0] > 0.5 * e[0] and b[1] > 0.5 * e[1]


def test_evidence_reuse_matches_prepared(ctx):
    """seg/tests.rs:229-266."""
    sp, n = opaque(ctx, row_attributes(4, 2.0, 0.7))
    sc = R.RenderScratch(ctx, n, 64, 64)
    mask = left_half_mask(ctx, 64, 64)
    fresh = G.Accumulators(ctx, n)
    it = R.prepare_isects(sp, sc, R.Camera(), 64, 64)
    G.accumulate_prepared(sc, it, fresh, mask)
    assert fresh.fg.read().any()
    it = R.prepare_isects(sp, sc, R.Camera(), 64, 64)
    R.finalize_rgb(sc, it)
    reused = G.Accumulators(ctx, n)
    G.accumulate_prepared(sc, it, reused, mask)
    np.testing.assert_array_equal(fresh.fg.read(), reused.fg.read())
    np.testing.assert_array_equal(fresh.bg.read(), reused.bg.read())


def test_scale_tracks_resolution_change(ctx):
    """seg/tests.rs:268-291."""
    sp, n = opaque(ctx, row_attributes(2, 2.0, 1.0))
    sc = R.RenderScratch(ctx, n, 64, 64)
    reused = G.Accumulators(ctx, n)
    G.render_responsibility(sp, sc, reused, R.Camera(), 64, 64)
    G.render_responsibility(sp, sc, reused, R.Camera(), 32, 32)
    fresh = G.Accumulators(ctx, n)
    G.render_responsibility(sp, sc, fresh, R.Camera(), 32, 32)
    np.testing.assert_array_equal(reused.bits.read(), fresh.bits.read())
    assert reused.scale == 2.0 ** 22


def test_beta_update_uses_evidence(ctx):
    """seg/tests.rs:293-310, bit-exact against the oracle's update."""
    _, _, ev = split_scene(ctx)
    st = G.BetaState.new_uniform(ctx, 2)
    st.update(ev)
    a, b = st.posteriors()
    assert a[0] > b[0] and b[1] > a[1]
    oa, ob = O.beta_update(np.ones(2), np.ones(2), ev.fg.read(), ev.bg.read(), ev.scale)
    np.testing.assert_array_equal(a, oa)
    np.testing.assert_array_equal(b, ob)


def test_eig_matches_oracle_f64(ctx):
    """EIG over a rendered responsibility map (C1 scene) and a random
    posterior: within 1e-3 nats of the oracle's f64 sum."""
    n = 65536
    attrs, sh, k = S.synthetic_scene(n, 0, seed=0)
    sp = R.Splats(ctx, attrs, sh)
    sc = R.RenderScratch(ctx, n, 256, 256)
    cam = R.Camera.look_at((0.0, 0.0, 3.0), (0.0, 0.0, 0.0))
    acc = G.Accumulators(ctx, n)
    G.render_responsibility(sp, sc, acc, cam, 256, 256)
    rng = np.random.default_rng(8)
    a = (1 + rng.random(n) * 30).astype(np.float32)
    b = (1 + rng.random(n) * 30).astype(np.float32)
    a[:100] = 5e5  # saturated splats
    st = G.BetaState(ctx, a, b)
    got = st.eig(acc)
    want = O.eig(a, b, acc.bits.read(), acc.scale)
    assert want > 0
    assert abs(got - want) <= 1e-3, (got, want)


def _entropy_exact(a, b):
    """Beta entropy from scipy's lgamma/digamma."""
    from scipy.special import betaln, digamma
    return (betaln(a, b) - (a - 1) * digamma(a) - (b - 1) * digamma(b)
            + (a + b - 2) * digamma(a + b))


def test_eig_reference_cases(ctx):
    """beta.rs:548-612: monotone and diminishing, invisible -> 0, fresh splat
    dH = H(1,1) - H(1.25,1.25)."""
    scale = 4096.0

    def eig1(a, b, eps):
        acc = G.Accumulators(ctx, len(a))
        acc.bits.write((np.asarray(eps, np.float32) * np.float32(scale)).astype(np.uint32))
        acc.scale = scale
        return G.BetaState(ctx, a, b).eig(acc)

    d = [eig1([aa], [bb], [0.7]) for aa, bb in [(1, 1), (3, 1), (50, 1)]]
    assert d[0] > d[1] > d[2] >= 0
    assert abs(eig1(np.ones(512), np.ones(512), np.zeros(512))) < 1e-12
    fresh = eig1([1.0], [1.0], [0.5])
    # The device entropy is accurate to ~1e-13; the reference's digamma
    # series (and the oracle restating it) carries ~1e-9.
    assert abs(fresh - (_entropy_exact(1, 1) - _entropy_exact(1.25, 1.25))) < 1e-11
    assert abs(fresh - (O.beta_entropy(1, 1) - O.beta_entropy(1.25, 1.25))) < 5e-9
    sat = eig1([5e5, 5e5, 1, 1], [5e5, 5e5, 1, 1], [0.5, 0, 0, 0])
    assert sat < 1e-2 < fresh


def test_entropy_drop_grid_against_scipy(ctx):
    """The device Beta entropy (the three arguments' shift products
    normalized and fused: four logs, three divisions) against scipy's
    lgamma / digamma: single-splat entropy drops over a grid of posteriors
    and evidence, within 1e-10 nats."""
    scale = 4096.0
    for a0, b0 in [(1.0, 1.0), (1.0, 40.0), (2.5, 7.25), (33.0, 33.0), (900.0, 3.0), (300.0, 700.0)]:
        for eps in (0.25, 3.0, 60.0):
            acc = G.Accumulators(ctx, 1)
            acc.bits.write(np.array([np.float32(eps) * np.float32(scale)], np.float32).astype(np.uint32))
            acc.scale = scale
            got = G.BetaState(ctx, np.array([a0], np.float32), np.array([b0], np.float32)).eig(acc)
            k = 1.0 + float(np.float32(eps)) / (a0 + b0)
            want = max(0.0, _entropy_exact(a0, b0) - _entropy_exact(a0 * k, b0 * k))
            assert abs(got - want) <= 1e-10, (a0, b0, eps, got, want)


def test_localize_matches_oracle(ctx):
    """Device views::localize against the oracle's f64 restatement on a 1M
    splat scene with a random posterior (f32 outputs, 1e-6 relative)."""
    n = 1 << 20
    attrs, sh, k = S.synthetic_scene(n, 0, seed=5)
    sp = R.Splats(ctx, attrs, sh)
    rng = np.random.default_rng(6)
    a = (1 + rng.random(n) * 4).astype(np.float32)
    b = (1 + rng.random(n) * 4).astype(np.float32)
    st = G.BetaState(ctx, a, b)
    got = G.localize(sp, st)
    want = O.localize(attrs, n, a, b)
    assert got is not None
    np.testing.assert_allclose(got[0], want[0], rtol=1e-6, atol=1e-6)
    assert abs(got[1] - want[1]) <= 1e-6 * want[1]
    none = G.localize(sp, G.BetaState(ctx, b * 0 + 1, b * 0 + 1))
    assert none is None


def test_segmenter_steps(ctx):
    """Segmenter::step (active.rs:133-183) over the device pipeline with a
    synthetic oracle (a centred disk): round 0 observes from the bootstrap
    camera with EIG 0; the disk's evidence moves the posterior; the next
    round localizes on the device and the batched and plain view scorings
    pick the same candidate with the same EIG."""
    n = 65536
    attrs, sh, k = S.synthetic_scene(n, 0, seed=0)
    sp = R.Splats(ctx, attrs, sh)
    cam0 = R.Camera.look_at((0.0, 0.0, 3.0), (0.0, 0.0, 0.0))
    res = 128
    yy, xx = np.mgrid[0:res, 0:res]
    disk = (((xx + 0.5 - res / 2) ** 2 + (yy + 0.5 - res / 2) ** 2) < (res / 4) ** 2).astype(np.uint8)
    seen = []

    def oracle(rgb, size, camera):
        assert rgb.size == size[0] * size[1] * 3
        seen.append(camera)
        return disk.reshape(-1)

    seg = G.Segmenter(sp, G.Config(resolution=res, candidates=8, iterations=2), cam0,
                      batched=True)
    r0 = seg.step(oracle)
    assert r0.camera is cam0 and r0.eig == 0.0 and r0.fg_pixels == int(disk.sum())
    a, b = seg.state.posteriors()
    assert (a >= 1).all() and (b >= 1).all() and (a > 1).any() and (b > 1).any()
    loc = G.localize(sp, seg.state)
    assert loc is not None
    cb, eb = seg.select_view(*loc)
    seg.batched = False
    cp, ep = seg.select_view(*loc)
    assert cb.position == cp.position and abs(eb - ep) <= 1e-6 * max(1.0, abs(ep))
    seg.batched = True
    r1 = seg.step(oracle)
    assert r1.camera.position == cb.position and r1.eig > 0.0
    assert len(seen) == 2
