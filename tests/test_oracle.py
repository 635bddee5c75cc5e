"""CPU checks that pin the oracle (oracle/sfo.c) before it is trusted as the
GPU parity reference: the reference's own known-answer tests, independent
library oracles (libm, scipy, numpy), closed forms and brute force."""
import math

import numpy as np
import pytest
from scipy import special

from oracle import oracle as O

from .fixtures import (GOLDEN_EPS, jitter, overdraw_attributes, row_attributes,
                       sample_opaque_attributes)


def default_view(w, h, tile_w=16, pos=(0, 0, 0), q=(0, 0, 0, 1), fov=(0.8, 0.8)):
    return O.camera_view(fov, pos, q, w, h, tile_w)


def ulp_dist(a, b):
    a = np.asarray(a, np.float32).view(np.int32).astype(np.int64)
    b = np.asarray(b, np.float32).view(np.int32).astype(np.int64)
    return np.abs(a - b)


# ------------------------------------------------------------------- math

def test_expf_matches_libm():
    rng = np.random.default_rng(1)
    xs = np.concatenate([rng.uniform(-104, 89, 20000), rng.uniform(-6, 6, 20000),
                         [0.0, -0.0, 1.0, -1.0, 88.0, -87.0, -100.0]]).astype(np.float32)
    got = np.array([O.expf(x) for x in xs], np.float32)
    want = np.exp(xs.astype(np.float64)).astype(np.float32)
    normal = want > np.float32(1.2e-38)
    assert ulp_dist(got[normal], want[normal]).max() <= 2
    # subnormal results: absolute error below one subnormal step pair
    assert np.all(np.abs(got[~normal].astype(np.float64) - np.exp(xs[~normal].astype(np.float64)))
                  <= 2.8e-45 * 2)
    assert O.expf(0.0) == 1.0
    assert O.expf(100.0) == np.inf and O.expf(-200.0) == 0.0


def test_logf_matches_libm():
    rng = np.random.default_rng(2)
    xs = np.concatenate([np.exp(rng.uniform(-80, 80, 20000)), rng.uniform(0.5, 2.0, 20000),
                         [1.0, 2.0, 0.5, 1e-40, 3.4e38]]).astype(np.float32)
    got = np.array([O.logf(x) for x in xs], np.float32)
    want = np.log(xs.astype(np.float64)).astype(np.float32)
    assert ulp_dist(got, want).max() <= 2
    assert O.logf(1.0) == 0.0
    assert math.isnan(O.logf(-1.0)) and O.logf(0.0) == -np.inf


# ----------------------------------------------------------------- camera

def test_camera_view_default_and_look_at():
    v = default_view(64, 64)
    a = v.as_array()
    np.testing.assert_array_equal(a[0:9], [1, 0, 0, 0, 1, 0, 0, 0, 1])
    np.testing.assert_array_equal(a[9:12], [0, 0, 0])
    f = np.float32(np.float32(32.0) / np.float32(math.tan(np.float32(0.4))))
    assert abs(a[12] - f) <= 1e-4 and a[19] == 4 and a[20] == 4
    # look_at from +Z back to the origin is pi about +Y (glam's axis choice)
    q = O.look_at((0, 0, 3), (0, 0, 0))
    assert abs(q[1] - 1.0) < 1e-7 and abs(q[3]) < 1e-7
    v = O.camera_view((1, 1), (0, 0, 3), q, 64, 64)
    z = np.array(v.rot2) @ np.zeros(3) + v.trans[2]
    assert abs(z - 3.0) < 1e-5  # the origin is 3 in front
    # generic look_at: forward axis points at the target
    for pos in [(1.0, 2.0, 3.0), (-2.0, 0.5, -1.0), (0.0, -3.0, 0.1)]:
        q = O.look_at(pos, (0, 0, 0))
        v = O.camera_view((1, 1), pos, q, 8, 8)
        fwd = np.array(v.rot2, np.float64)
        d = -np.array(pos, np.float64)
        assert np.dot(fwd, d / np.linalg.norm(d)) > 1 - 1e-5


def test_bits_for_table():
    """sort.rs:420-431."""
    for m, b in [(0, 0), (1, 0), (2, 1), (3, 2), (255, 8), (256, 8), (257, 9), (65536, 16),
                 (2**32 - 1, 32)]:
        assert O.bits_for(m) == b


def test_fractional_bits():
    """seg/mod.rs:31-34 closed form: 32 - log2(next_pow2(w*h))."""
    for w, h, fb in [(70, 70, 19), (64, 64, 20), (512, 512, 14), (1920, 1080, 11),
                     (256, 256, 16), (1024, 1024, 12)]:
        assert O.fractional_bits(w, h) == fb


# ---------------------------------------------------------------- project

def test_project_kat():
    """project.rs:399-446."""
    attrs = np.array([0, 0, 1, 1, 0, 0, 0, 0, 0, 0, 5], np.float32)
    pr = O.project(default_view(64, 64), attrs, np.zeros(3, np.float32), 1, 1)
    np.testing.assert_array_equal(pr["counters"], [16, 1])
    np.testing.assert_array_equal(pr["tile_counts"], [16])
    np.testing.assert_array_equal(pr["bbox"], [0, 4 | 4 << 16])


def test_project_closed_form_isotropic():
    """docs/01_project.md: an isotropic splat of scale s at depth z on the axis
    has cov2d = (f s / z)^2 + 0.3 on the diagonal, 0 off it; conic = 1/diag;
    mean at the image centre; ext = sqrt(2 ln(o/cutoff) cov)."""
    for s, z, o_logit in [(0.1, 2.0, 2.0), (0.02, 3.0, 0.0), (0.5, 5.0, 4.0)]:
        a = np.array([0, 0, z, 1, 0, 0, 0, math.log(s), math.log(s), math.log(s), o_logit],
                     np.float32)
        v = default_view(256, 256)
        pr = O.project(v, a, np.zeros(3, np.float32), 1, 1)
        f = v.focal[0]
        cov = (f * s / z) ** 2 + 0.3
        P = pr["projected"]
        assert P[0] == 128 and P[1] == 128
        assert abs(P[2] - 1 / cov) <= 1e-5 / cov and abs(P[3]) < 1e-9 and abs(P[4] - 1 / cov) <= 1e-5 / cov
        o = 1 / (1 + math.exp(-o_logit))
        assert abs(P[8] - o) < 1e-6
        ext = math.sqrt(2 * math.log(o / (10 / 255)) * cov)
        lo = max(0, int((128 - ext) / 16))
        hi = min(16, int((128 + ext) / 16 + 1))
        assert pr["bbox"][0] == lo | lo << 16 and pr["bbox"][1] == hi | hi << 16


def test_project_culls():
    """Opacity below 10/255 and depth <= 0.1 are culled (project.rs:267-276)."""
    a = sample_opaque_attributes(3, lambda i: [(0, 0, 1), (0, 0, 0.05), (0, 0, -2)][i])
    a = a.reshape(11, 3)
    a[10, 0] = -4.0  # sigmoid(-4) = 0.018 < 10/255
    pr = O.project(default_view(32, 32), a.reshape(-1), np.zeros(9, np.float32), 3, 1)
    np.testing.assert_array_equal(pr["visible"], [0, 0, 0])
    np.testing.assert_array_equal(pr["counters"], [0, 0])


def test_project_invariants_random_scene():
    from splatfield_amd import scene as S
    n = 4096
    attrs, sh, k = S.synthetic_scene(n, 3, seed=5)
    q = O.look_at((0, 0, 3), (0, 0, 0))
    v = O.camera_view((1.0, 1.0), (0, 0, 3), q, 128, 96)
    pr = O.project(v, attrs, sh, n, k)
    vis = pr["visible"].astype(bool)
    assert pr["counters"][1] == vis.sum()
    assert pr["counters"][0] == pr["tile_counts"][vis].sum()
    bb = pr["bbox"].reshape(n, 2)[vis]
    mnx, mny, mxx, mxy = bb[:, 0] & 0xFFFF, bb[:, 0] >> 16, bb[:, 1] & 0xFFFF, bb[:, 1] >> 16
    assert np.all(mnx <= mxx) and np.all(mxx <= 8) and np.all(mxy <= 6)
    np.testing.assert_array_equal((mxx - mnx) * (mxy - mny), pr["tile_counts"][vis])
    # depth key is the f32 bits of camera z: monotone in z
    z = pr["depth_keys"][vis].view(np.float32)
    assert np.all(z > 0.1)


def _real_sh_basis(x, y, z):
    """Textbook real SH basis (orthonormal, Condon-Shortley) in the INRIA
    3DGS ordering for bands 0-2, with the sign conventions of
    project.rs:351-355 (band 1: -y, z, -x)."""
    c0 = 0.28209479177387814
    c1 = 0.4886025119029199
    c2 = [1.0925484305920792, -1.0925484305920792, 0.31539156525252005,
          -1.0925484305920792, 0.5462742152960396]
    return [c0, -c1 * y, c1 * z, -c1 * x,
            c2[0] * x * y, c2[1] * y * z, c2[2] * (2 * z * z - x * x - y * y),
            c2[3] * x * z, c2[4] * (x * x - y * y)]


def test_sh_bands_0_to_2_match_textbook_basis():
    """project.rs:342-367 vs the textbook real SH basis: each coefficient in
    turn set to 1 must reproduce its basis function at the view direction."""
    rng = np.random.default_rng(7)
    for _ in range(5):
        pos = rng.normal(size=3)
        pos = (pos / np.linalg.norm(pos) * 0.5).astype(np.float32)
        d = pos - np.array([0, 0, -3], np.float32)
        d = d / np.linalg.norm(d)
        basis = _real_sh_basis(*d)
        for k in range(9):
            a = sample_opaque_attributes(1, lambda i: pos)
            sh = np.zeros(27, np.float32)
            sh[k * 3 + 0] = 1.0  # red channel of coefficient k
            q = np.array([0, 0, 0, 1], np.float32)
            v = O.camera_view((1.0, 1.0), (0, 0, -3), q, 32, 32)
            pr = O.project(v, a, sh, 1, 9)
            assert abs((pr["projected"][5] - 0.5) - basis[k]) < 2e-6, k


# Band-3 normalisations of the real SH basis, from their closed forms.
_K3_33 = math.sqrt(35.0 / (2.0 * math.pi)) / 4.0    # x(x^2-3y^2), y(3x^2-y^2)
_K3_32 = math.sqrt(105.0 / math.pi) / 2.0           # xyz
_K3_31 = math.sqrt(21.0 / (2.0 * math.pi)) / 4.0    # x(5z^2-1), y(5z^2-1)
_K3_30 = math.sqrt(7.0 / math.pi) / 4.0             # z(5z^2-3)
_K3_2z = math.sqrt(105.0 / math.pi) / 4.0           # z(x^2-y^2)


def _band3_as_written(x, y, z):
    """Band 3 of project.rs:362-384 by coefficient slot 9..15, derived by
    hand.  Slots 11..15 are the textbook real SH functions (INRIA sign
    convention) at permuted slots: y(5z^2-1), x(5z^2-1), z(5z^2-3), xyz,
    z(x^2-y^2).  Slots 9 and 10 are not: with Rust precedence
    b9 = (-c*x)*(2xy) + y*(x^2-y^2) and b10 = (-c*x)*(x^2-y^2) - y*(2xy),
    which expand to b9 = (1-2c)x^2y - y^3 and b10 = -c x^3 + (c-2)xy^2."""
    c = _K3_33
    return {9: (1 - 2 * c) * x * x * y - y ** 3,
            10: -c * x ** 3 + (c - 2) * x * y * y,
            11: -_K3_31 * y * (5 * z * z - 1),
            12: -_K3_31 * x * (5 * z * z - 1),
            13: _K3_30 * z * (5 * z * z - 3),
            14: _K3_32 * x * y * z,
            15: _K3_2z * z * (x * x - y * y)}


def _sh3_red(direction, k):
    """Red colour - 0.5 of one splat seen along `direction` whose only
    non-zero SH coefficient is coefficient k's red channel (slot 3k)."""
    d = np.asarray(direction, np.float64)
    pos = (2.0 * d / np.linalg.norm(d)).astype(np.float32)
    q = O.look_at((0, 0, 0), pos)
    v = O.camera_view((1.0, 1.0), (0, 0, 0), q, 32, 32)
    a = sample_opaque_attributes(1, lambda i: pos)
    sh = np.zeros(48, np.float32)
    sh[3 * k] = 1.0
    pr = O.project(v, a, sh, 1, 16)
    assert pr["counters"][1] == 1
    return float(pr["projected"][5]) - 0.5


def test_sh_band3_matches_textbook_basis_and_hand_values():
    """project.rs:362-384 (band 3): b11..b15 against the textbook real SH
    functions with normalisations from their closed forms, at random unit
    directions; b9/b10 on the axes, where they reduce to -y^3 (b9 = -1 along
    +y, 0 along +-x) and -c x^3 (b10 = -c along +x, 0 along +-y), and at
    random directions against their hand expansion.  A slot swap, a sign or
    the bracketed -c*(...) form of b9/b10 fails it."""
    rng = np.random.default_rng(11)
    dirs = [rng.normal(size=3) for _ in range(6)]
    dirs = [d / np.linalg.norm(d) for d in dirs]
    for d in dirs:
        want = _band3_as_written(*d)
        for k in range(9, 16):
            got = _sh3_red(d, k)
            assert abs(got - want[k]) < 3e-6, (k, d, got, want[k])
    c = np.float32(_K3_33)
    axis = {(1, 0, 0): {9: 0.0, 10: -c}, (-1, 0, 0): {9: 0.0, 10: c},
            (0, 1, 0): {9: -1.0, 10: 0.0}, (0, -1, 0): {9: 1.0, 10: 0.0}}
    for d, vals in axis.items():
        for k, want in vals.items():
            assert abs(_sh3_red(d, k) - float(want)) < 1e-6, (d, k)
    # the test tells the as-written b9/b10 from the bracketed form
    # -c*(x*(2xy) + y*(x^2-y^2)) and -c*(x*(x^2-y^2) - y*(2xy)) at these
    # directions, so a precedence slip in the oracle could not pass it
    for d in dirs:
        x, y, _ = d
        bracket9 = -_K3_33 * (x * 2 * x * y + y * (x * x - y * y))
        bracket10 = -_K3_33 * (x * (x * x - y * y) - y * 2 * x * y)
        w = _band3_as_written(*d)
        assert max(abs(bracket9 - w[9]), abs(bracket10 - w[10])) > 1e-3


# ------------------------------------------------------------- sort / scan

@pytest.mark.parametrize("n", [1000, 5000, 200000])
def test_argsort_stable_vs_numpy(n):
    """sort.rs:433-453: stable argsort of keys < 37 at 8 bits."""
    rng = np.random.default_rng(n)
    keys = rng.integers(0, 37, n).astype(np.uint32)
    ok, ov = O.argsort_u32(keys, np.arange(n, dtype=np.uint32), 8)
    np.testing.assert_array_equal(ov, np.argsort(keys, kind="stable"))
    np.testing.assert_array_equal(ok, np.sort(keys))


def test_argsort_32bit_and_u64():
    rng = np.random.default_rng(4)
    keys = rng.integers(0, 2**32, 50000, dtype=np.uint64).astype(np.uint32)
    ok, ov = O.argsort_u32(keys, np.arange(keys.size, dtype=np.uint32), 32)
    np.testing.assert_array_equal(ov, np.argsort(keys, kind="stable"))
    k64 = rng.integers(0, 2**63, 50000, dtype=np.uint64)
    k64[::7] = k64[3]  # ties
    ok, ov = O.argsort_u64(k64, np.arange(k64.size, dtype=np.uint32), 64)
    np.testing.assert_array_equal(ov, np.argsort(k64, kind="stable"))


@pytest.mark.parametrize("n", [1, 5, 1024, 1025, 70000, 100000])
def test_scan_gather_closed_form(n):
    """scan.rs:283-331: offsets = exclusive cumsum of counts[gids]."""
    rng = np.random.default_rng(n)
    counts = (np.arange(n) % 4).astype(np.uint32) if n in (1025, 70000) else \
        rng.integers(0, 20, n).astype(np.uint32)
    gids = rng.permutation(n).astype(np.uint32)
    got = O.scan_gather(gids, counts)
    want = np.concatenate([[0], np.cumsum(counts[gids])[:-1]]).astype(np.uint32)
    np.testing.assert_array_equal(got, want)


def test_lower_bound_matches_partition_point():
    """render.rs:824-850."""
    ids = np.array([0, 0, 1, 2, 2, 2, 5, 5], np.uint32)
    for t in range(8):
        assert O.lower_bound(ids, t) == int(np.searchsorted(ids, t, side="left"))
    r = O.tile_ranges(ids, 8)
    for t in range(8):
        assert r[2 * t + 1] - r[2 * t] == int(np.sum(ids == t))


def test_map_isects_brute_force():
    """map_isects emits exactly each visible splat's bbox tiles in (rank,
    row-major tile) order; the stable tile sort yields tile-major /
    depth-minor order (render.rs:342-352)."""
    from splatfield_amd import scene as S
    n = 3000
    attrs, sh, k = S.synthetic_scene(n, 0, seed=9)
    q = O.look_at((0, 0, 3), (0, 0, 0))
    v = O.camera_view((1.0, 1.0), (0, 0, 3), q, 96, 80)
    pr = O.prepare(v, attrs, sh, n, k, 96, 80)
    rank = np.empty(n, np.int64)
    rank[pr["depth_order"]] = np.arange(len(pr["depth_order"]))
    want = []
    bb = pr["bbox"].reshape(n, 2)
    for r, g in enumerate(pr["depth_order"]):
        lo, hi = bb[g]
        for ty in range(lo >> 16, hi >> 16):
            for tx in range(lo & 0xFFFF, hi & 0xFFFF):
                want.append((tx + ty * pr["tiles_x"], r, g))
    want.sort(key=lambda t: (t[0], t[1]))
    np.testing.assert_array_equal(pr["tile_ids"], [t[0] for t in want])
    np.testing.assert_array_equal(pr["gaussian_ids"], [t[2] for t in want])
    # depth order: non-decreasing keys, ties by splat index
    keys = pr["depth_keys"][pr["depth_order"]]
    assert np.all(keys[1:] >= keys[:-1])
    ties = keys[1:] == keys[:-1]
    assert np.all(pr["depth_order"][1:][ties] > pr["depth_order"][:-1][ties])


# ----------------------------------------------------------------- raster

def _render_default(attrs, n, w, h, mode=O.MODE_RGB, scale=0.0, mask=None, wpr=0):
    return O.render(default_view(w, h), attrs, np.zeros(3 * n, np.float32), n, 1, w, h, 16,
                    mode, scale, mask, wpr)


def test_golden_center():
    """render.rs:671-710."""
    n = 5
    s2 = np.float32(np.sin(np.float32(np.pi / 4)))
    c2 = np.float32(np.cos(np.float32(np.pi / 4)))
    for attrs, pos, q in [
        (sample_opaque_attributes(n, lambda i: (0, 0, 1 + 0.5 * i)), (0, 0, 0), (0, 0, 0, 1)),
        (sample_opaque_attributes(n, lambda i: (0, 0, 0)), (-2, 0, 0), (0, s2, 0, c2)),
    ]:
        r = O.render(default_view(32, 32, 16, pos, q), attrs, np.zeros(15, np.float32), n, 1,
                     32, 32)
        bm = r["bitmap"].reshape(32, -1)
        c = int(bm[16, 16])
        assert c >> 24 >= 254 and int(bm[0, 0]) >> 24 == 0
        for s in (0, 8, 16):
            assert abs(((c >> s) & 255) - 127) <= 1


def test_eps_golden_words():
    """seg/tests.rs:186-203: GOLDEN from the reference's device exp; the
    oracle's exp differs in the last ulp, so words match within 2e-7
    relative (observed: <= 157 LSB on 4.1e8), and zeros match exactly."""
    n = 40
    r = _render_default(overdraw_attributes(), n, 70, 70, O.MODE_TOTAL, float(2 ** 19))
    got = r["resp"].astype(np.int64)
    gold = np.array(GOLDEN_EPS, np.int64)
    assert np.all((got == 0) == (gold == 0))
    assert np.all(np.abs(got - gold) <= np.maximum(8, gold * 5e-7))


def test_responsibility_conserves_pixel_weights():
    """seg/tests.rs:112-126."""
    n = 5
    a = sample_opaque_attributes(n, lambda i: (0, 0, 1 + 0.5 * i))
    rgb = _render_default(a, n, 64, 64)
    alpha_sum = (rgb["bitmap"].reshape(64, -1)[:, :64] >> 24).astype(np.float64).sum() / 255
    eps = _render_default(a, n, 64, 64, O.MODE_TOTAL, float(2 ** 20))
    eps_sum = eps["resp"].astype(np.float64).sum() / 2 ** 20
    assert abs(alpha_sum - eps_sum) < 64 * 64 * 3 / 255


def test_responsibility_orders_depth_and_culls():
    """seg/tests.rs:128-150."""
    n = 6
    a = sample_opaque_attributes(n, lambda i: (0, 0, -3 if i == 5 else 1 + 0.5 * i))
    eps = _render_default(a, n, 64, 64, O.MODE_TOTAL, float(2 ** 20))["resp"]
    assert all(eps[k] > eps[k + 1] for k in range(n - 2))
    assert eps[5] == 0


def test_evidence_split_and_update():
    """seg/tests.rs:205-227, 293-310: fg + bg == total bit for bit; the
    update moves a/b by fg/scale, bg/scale."""
    n = 2
    a = row_attributes(n, 2.0, 1.0)
    mask = np.zeros((64, 64), np.uint8)
    mask[:, :32] = 1
    words = np.zeros(2 * 64, np.uint32)
    for y in range(64):
        for x in range(32):
            words[y * 2 + x // 32] |= 1 << (x % 32)
    scale = float(2 ** 20)
    tot = _render_default(a, n, 64, 64, O.MODE_TOTAL, scale)["resp"]
    ev = _render_default(a, n, 64, 64, O.MODE_EVIDENCE, scale, words, 2)
    np.testing.assert_array_equal(ev["fg"] + ev["bg"], tot)
    assert ev["bg"][0] == 0 and ev["fg"][1] == 0
    aa, bb = O.beta_update(np.ones(2), np.ones(2), ev["fg"], ev["bg"], scale)
    assert aa[0] > bb[0] and bb[1] > aa[1]
    assert abs(aa[0] - (1 + ev["fg"][0] / scale)) < 1e-5


def test_tile_width_changes_only_low_alpha_tails():
    """Tile size changes which splats a tile lists (the bbox uses the 10/255
    cutoff, the compositor 1/255), so only alpha < 10/255 tails differ."""
    n = 5
    a = sample_opaque_attributes(n, lambda i: (0, 0, 1 + 0.5 * i))
    imgs = []
    for tw in (8, 16, 32):
        r = O.render(default_view(64, 64, tw), a, np.zeros(15, np.float32), n, 1, 64, 64, tw)
        imgs.append(r["rgba"])
    assert np.max(np.abs(imgs[0] - imgs[1])) <= 10 / 255
    assert np.max(np.abs(imgs[2] - imgs[1])) <= 10 / 255
    assert np.abs(imgs[2] - imgs[1]).mean() < 1e-3


# -------------------------------------------------------------------- seg

def test_entropy_vs_scipy():
    """beta.rs:55-96 vs scipy's betaln / digamma (an independent library)."""
    rng = np.random.default_rng(3)
    pairs = [(1.0, 1.0), (1.5, 1.5), (2.0, 5.0), (9.0, 1.0), (30.0, 30.0), (100.0, 3.0)]
    pairs += [tuple(np.exp(rng.uniform(0, 14, 2))) for _ in range(256)]
    for a, b in pairs:
        want = (special.betaln(a, b) - (a - 1) * special.digamma(a) - (b - 1) * special.digamma(b)
                + (a + b - 2) * special.digamma(a + b))
        got = O.beta_entropy(a, b)
        assert abs(got - want) <= 1e-7 + 1e-10 * abs(want), (a, b, got, want)
        assert abs(O.lgamma(a) - special.gammaln(a)) <= 1e-10 * max(1, abs(special.gammaln(a)))
        assert abs(O.digamma(a) - special.digamma(a)) <= 5e-9  # series truncated at x^-6, x >= 6
    assert abs(O.beta_entropy(1.0, 1.0)) < 1e-12


def test_eig_mirror_and_fresh_splat():
    """beta.rs:501-547, 587-612: EIG equals the sum of per-splat entropy drops;
    a fresh splat seen at eps = 0.5 gains H(1,1) - H(1.25,1.25)."""
    rng = np.random.default_rng(6)
    n = 1000
    a = (1 + rng.random(n) * 20).astype(np.float32)
    b = (1 + rng.random(n) * 20).astype(np.float32)
    eps = np.where(np.arange(n) % 2 == 0, rng.random(n) * 0.5, 0).astype(np.float32)
    scale = 4096.0
    q = (eps * np.float32(scale)).astype(np.uint32)
    got = O.eig(a, b, q, scale)
    want = 0.0
    for i in range(n):
        e = np.float32(q[i]) / np.float32(scale)
        if e > 0:
            k = 1 + float(e) / (float(a[i]) + float(b[i]))
            ai, bi = float(a[i]), float(b[i])
            d = O.beta_entropy(ai, bi) - O.beta_entropy(ai * k, bi * k)
            want += max(d, 0.0)
    assert abs(got - want) < 1e-9
    q2 = (2 * eps * np.float32(scale)).astype(np.uint32)
    assert O.eig(a, b, q2, scale) > got > 0
    fresh = O.eig([1.0], [1.0], [int(0.5 * 4096)], 4096.0)
    assert abs(fresh - (O.beta_entropy(1, 1) - O.beta_entropy(1.25, 1.25))) < 1e-12
    assert O.eig(np.ones(512), np.ones(512), np.zeros(512, np.uint32), 4096.0) == 0.0
    d = [O.eig([aa], [1.0], [int(0.7 * 4096)], 4096.0) for aa in (1.0, 3.0, 50.0)]
    assert d[0] > d[1] > d[2] > 0


def test_candidate_sphere():
    """views.rs:279-310: distance, facing, decreasing z."""
    from splatfield_amd import scene as S
    dirs = S.fibonacci_directions(20)
    assert np.allclose(np.linalg.norm(dirs, axis=1), 1, atol=1e-6)
    assert np.all(np.diff(dirs[:, 2]) < 0)


def test_jitter_fixture():
    assert np.all(np.abs(np.array([jitter(i) for i in range(100)])) <= 0.5)


def test_localize_closed_forms():
    """views.rs:32-67: equal-weight foreground on a sphere of radius 2 about
    (1, -1, 0.5) gives that centre and radius 2; background and ties (a == b)
    are ignored; no foreground gives None; weights a/(a+b) pull the centre
    toward the heavier splat."""
    rng = np.random.default_rng(3)
    n = 4000
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    c0 = np.array([1.0, -1.0, 0.5])
    # antipodal pairs: the centroid is exactly c0
    pts = np.concatenate([c0 + 2.0 * d[: n // 2], c0 - 2.0 * d[: n // 2]])
    attrs = np.zeros((11, n), np.float32)
    attrs[:3] = pts.T
    a = np.full(n, 3.0, np.float32)
    b = np.ones(n, np.float32)
    c, r = O.localize(attrs, n, a, b)
    np.testing.assert_allclose(c, c0, atol=2e-6)
    assert abs(r - 2.0) < 2e-6
    b2 = b.copy()
    b2[n // 4:] = 3.0  # ties (a == b) are background
    c2, _ = O.localize(attrs, n, a, b2)
    sel = pts[: n // 4]
    np.testing.assert_allclose(c2, sel.mean(0), atol=1e-6)
    assert O.localize(attrs, n, b, a) is None
    attrs3 = np.zeros((11, 2), np.float32)
    attrs3[0] = [0.0, 1.0]
    c3, r3 = O.localize(attrs3, 2, np.array([3.0, 1.5], np.float32), np.array([1.0, 1.0], np.float32))
    w0, w1 = 0.75, 0.6
    assert abs(c3[0] - w1 / (w0 + w1)) < 1e-7
    assert abs(r3 - 2 * w0 * w1 / (w0 + w1) ** 2) < 1e-7


# ------------------------------------------------------------------ paint

def test_paint_posterior_pins():
    """paint.rs:96-109: the posterior render is sum(m_i alpha_i T_i) with no
    +0.5.  m = 1 gives 1 - T (the alpha channel), m = 0 gives black, halving
    every m halves every channel exactly (a power of two), and T is the
    colour render's: a paint that touched a non-colour field, added the
    offset or missed a channel fails one of these."""
    n = 40
    attrs = overdraw_attributes()
    sh = np.zeros(3 * n, np.float32)
    view = default_view(70, 70)
    plain = O.render(view, attrs, sh, n, 1, 70, 70)
    ones, zeros = np.ones(n, np.float32), np.zeros(n, np.float32)
    white = O.render_posterior(view, attrs, sh, n, 1, 70, 70, ones, zeros)
    black = O.render_posterior(view, attrs, sh, n, 1, 70, 70, zeros, ones)
    rng = np.random.default_rng(3)
    a = rng.integers(1, 9, n).astype(np.float32)
    half = O.render_posterior(view, attrs, sh, n, 1, 70, 70, a, a)
    quarter = O.render_posterior(view, attrs, sh, n, 1, 70, 70, a, 3 * a)
    for r in (white, black, half, quarter):
        np.testing.assert_array_equal(r["final_t"], plain["final_t"])
        px = r["rgba"].reshape(-1, 4)
        np.testing.assert_array_equal(px[:, 3], plain["rgba"].reshape(-1, 4)[:, 3])
        assert np.all(px[:, 0] == px[:, 1]) and np.all(px[:, 1] == px[:, 2])
    w4 = white["rgba"].reshape(-1, 4)
    assert w4[:, 3].max() > 0.9
    np.testing.assert_allclose(w4[:, 0], w4[:, 3], atol=2e-6)
    assert not black["rgba"].reshape(-1, 4)[:, :3].any()
    np.testing.assert_array_equal(quarter["rgba"].reshape(-1, 4)[:, :3] * 2,
                                  half["rgba"].reshape(-1, 4)[:, :3])
    np.testing.assert_array_equal(half["rgba"].reshape(-1, 4)[:, :3] * 2, w4[:, :3])


def test_tint_overwrites_foreground_dc_only():
    """tint_colors (paint.rs:113-122): DC planes c*n + i take the colour
    where a > b (ties are background); every other float is unchanged."""
    n = 1000
    rng = np.random.default_rng(4)
    sh = rng.normal(size=3 * 16 * n).astype(np.float32)
    a = rng.integers(0, 4, n).astype(np.float32)
    b = rng.integers(0, 4, n).astype(np.float32)
    out = O.tint(sh, a, b, (0.5, -1.0, 2.0))
    fg = a > b
    assert 0 < fg.sum() < n and np.any(a == b)
    dc = out[:3 * n].reshape(3, n)
    for c, v in enumerate((0.5, -1.0, 2.0)):
        assert np.all(dc[c, fg] == np.float32(v))
        np.testing.assert_array_equal(dc[c, ~fg], sh[c * n:(c + 1) * n][~fg])
    np.testing.assert_array_equal(out[3 * n:], sh[3 * n:])


# -------------------------------------------------------------------- PLY

def _synthetic_ply(n, extra):
    """ply.rs tests' synthetic_ply: the 14 base properties then extra, value
    (i*31 + j)*0.25 - 8 at row i, column j."""
    base = ["x", "y", "z", "rot_0", "rot_1", "rot_2", "rot_3", "scale_0", "scale_1", "scale_2",
            "opacity", "f_dc_0", "f_dc_1", "f_dc_2"]
    props = base + list(extra)
    head = "ply\nformat binary_little_endian 1.0\nelement vertex %d\n" % n
    head += "".join(f"property float {p}\n" for p in props) + "end_header\n"
    st = len(props)
    rows = np.array([[(i * 31 + j) * 0.25 - 8.0 for j in range(st)] for i in range(n)], "<f4")
    return head.encode() + rows.tobytes(), rows


def _both_parsers(data):
    """(oracle result or error text, C header parser error text or layout)."""
    from splatfield_amd import render as R
    try:
        o = O.parse_ply(data)
    except ValueError as e:
        o = str(e)
    try:
        c = R.parse_ply_header(data)
    except ValueError as e:
        c = str(e)
    return o, c


def test_ply_layout_kat():
    """ply.rs test_parse_ply_layout: field-major planes, DC planes, f_rest
    sorted by number and moved channel-major -> interleaved; the rotation
    within an ulp of the f64 normalisation."""
    data, rows = _synthetic_ply(3, ["f_rest_0", "f_rest_2", "f_rest_1"])
    (attrs, sh), lay = _both_parsers(data)
    assert attrs.size == 33 and sh.size == 18
    assert (lay.n, lay.stride, lay.n_rest, lay.sh_planes) == (3, 17, 3, 6)
    assert list(lay.rest)[:3] == [14, 16, 15]
    assert lay.data_offset == data.index(b"end_header") + 11
    a = attrs.reshape(11, 3)
    for k, col in ((0, 0), (1, 1), (2, 2), (7, 7), (8, 8), (9, 9), (10, 10)):
        np.testing.assert_array_equal(a[k], rows[:, col])
    s = sh.reshape(6, 3)
    for c in range(3):
        np.testing.assert_array_equal(s[c], rows[:, 11 + c])
    np.testing.assert_array_equal(s[3], rows[:, 14])  # f_rest_0 -> coefficient 1, red
    np.testing.assert_array_equal(s[4], rows[:, 16])  # f_rest_1 -> green
    np.testing.assert_array_equal(s[5], rows[:, 15])  # f_rest_2 -> blue
    q = rows[:, [3, 4, 5, 6]].astype(np.float64)
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    assert np.max(ulp_dist(a[3:7].T, q.astype(np.float32))) <= 1


@pytest.mark.parametrize("case,msg", [
    ("missing", "Missing property: f_dc_2"), ("empty", "PLY contains no vertices"),
    ("ascii", "Unsupported PLY format: ascii"),
    ("double", "Unsupported property type double: f_rest_0"),
    ("short", "failed to read vertex 2 of 3")])
def test_ply_rejections(case, msg):
    """ply.rs tests: missing property, empty, other formats, non-float
    properties; and a file cut short (read_exact).  Oracle and the C header
    parser give the same message."""
    data, _ = _synthetic_ply(3, [])
    if case == "missing":
        data = data.replace(b"property float f_dc_2\n", b"")
    elif case == "empty":
        data = data.replace(b"element vertex 3", b"element vertex 0")
    elif case == "ascii":
        data = data.replace(b"binary_little_endian", b"ascii")
    elif case == "double":
        data = data.replace(b"end_header", b"property double f_rest_0\nend_header")
    else:
        data = data[:-10]
    o, c = _both_parsers(data)
    assert o == msg and c == msg, (o, c)


@pytest.mark.parametrize("k_per_ch", [1, 3, 16])
def test_ply_write_parse_roundtrip(k_per_ch):
    """ply.rs test_write_ply_roundtrips_parse: identity rotations, every
    plane back exactly."""
    n = 3
    attrs = np.array([((i * 7 % 23) - 11.0) / 8.0 for i in range(11 * n)], np.float32)
    attrs[3 * n:4 * n] = 1.0
    attrs[4 * n:7 * n] = 0.0
    sh = (np.arange(n * k_per_ch * 3, dtype=np.float32) * 0.25 - 9.0)
    a2, s2 = O.parse_ply(O.write_ply(attrs, sh, n))
    np.testing.assert_array_equal(a2, attrs)
    np.testing.assert_array_equal(s2, sh)


@pytest.mark.parametrize("deg", [0, 1, 3])
def test_product_ply_writer_matches_oracle(deg):
    """scene.write_ply (the library's CpuSplats::write_ply) writes the
    oracle writer's bytes, and the oracle parses them back exactly."""
    from splatfield_amd import scene as S
    n = 257
    attrs, sh, _ = S.synthetic_scene(n, deg, seed=9)
    data = S.write_ply(attrs, sh, n)
    assert data == O.write_ply(attrs, sh, n)
    a2, s2 = O.parse_ply(data)
    np.testing.assert_array_equal(s2, sh)
    np.testing.assert_array_equal(a2[:3 * n], attrs[:3 * n])
    np.testing.assert_array_equal(a2[7 * n:], attrs[7 * n:])


# -------------------------------------------------------------------- SOG

def test_sog_reference_unit_tests():
    """sog.rs tests: palette_offset row width, unit quaternions for every
    tag, inv_log / logit round trip."""
    assert O.sog_palette_offset(20, 128, 8) == (32, 1)
    for tag in range(252, 256):
        q = np.array(O.sog_unpack_quat(200, 10, 77, tag), np.float32)
        assert abs(float((q * q).sum()) - 1.0) < 1e-5
        assert q[tag - 252] == q.max() or tag == 255  # the rebuilt component sits at tag - 252
    info = dict(n=5, mins=[-3, -3, -3], maxs=[3, 3, 3], sh_per_ch=1)
    for v in (-3.0, -0.5, 0.0, 0.5, 3.0):
        y = np.float32(1.0) / (np.float32(1.0) + np.exp(np.float32(-v)))
        e = np.clip(np.float32(y), np.float32(1e-6), np.float32(1) - np.float32(1e-6))
        assert abs(float(np.log(np.float64(e / (np.float32(1) - e)))) - v) < 1e-4


def test_sog_host_reader_and_oracle():
    """WebP sheets survive the archive exactly; the oracle decode's means hit
    the closed forms at q16 = 0 and 65535, and every rejection carries the
    reference's message."""
    from splatfield_amd import scene as S
    from .fixtures import make_sog
    data, sheets = make_sog(300, 2, seed=1)
    info, got, cw, cbs = S.read_sog(data)
    for k, v in sheets.items():
        np.testing.assert_array_equal(got[k], v)
    assert info["sh_per_ch"] == 9 and cw == 16 * 8 + 3 and cbs.size == 768
    got["lo"][:4] = 0; got["hi"][:4] = 0          # splat 0: t = 0 on x, y, z
    got["lo"][4:8] = 255; got["hi"][4:8] = 255    # splat 1: t = 1
    attrs, sh = O.sog_decode(info, got, cw, cbs)
    a = attrs.reshape(11, 300)
    for k in range(3):
        assert a[k, 0] == np.float32(np.sign(info["mins"][k]) * np.expm1(abs(info["mins"][k])))
        assert a[k, 1] == np.float32(np.sign(info["maxs"][k]) * np.expm1(abs(info["maxs"][k])))
    q = a[3:7].astype(np.float64)
    slot = got["quats"].reshape(-1, 4)[:300, 3].astype(int) - 252
    rebuilt = q[slot, np.arange(300)]
    norm = (q ** 2).sum(0)
    # Unit norm wherever the three stored components fit in the unit ball;
    # random texels past it rebuild the fourth as 0 (the max(0) clamp).
    inside = rebuilt > 0
    assert inside.mean() > 0.8
    np.testing.assert_allclose(norm[inside], 1.0, atol=1e-5)
    assert np.all(rebuilt[~inside] == 0) and np.all(norm[~inside] >= 1.0 - 1e-6)
    assert np.all(np.isfinite(sh))
    bad, _ = make_sog(64, 1, seed=2, bad_tag_at=9)
    with pytest.raises(ValueError, match="quats: invalid rotation tag 17"):
        O.sog_decode(*S.read_sog(bad))
    with pytest.raises(ValueError, match="not a SOG archive"):
        S.read_sog(b"not a zip")
    with pytest.raises(ValueError, match="SOG contains no splats"):
        S.read_sog(_sog_with_meta(dict(_SOG_META, count=0)))


_SOG_META = {"count": 4, "means": {"mins": [0, 0, 0], "maxs": [1, 1, 1], "files": ["a", "b"]},
             "scales": {"codebook": [0.0] * 256, "files": ["c"]}, "quats": {"files": ["d"]},
             "sh0": {"codebook": [0.0] * 256, "files": ["e"]}}


def _sog_with_meta(meta):
    import io
    import json
    import zipfile
    buf = io.BytesIO()
    with zipfile.ZipFile(buf, "w") as z:
        z.writestr("meta.json", json.dumps(meta))
    return buf.getvalue()


def _meta_edit(path, value):
    import copy
    m = copy.deepcopy(_SOG_META)
    d = m
    for k in path[:-1]:
        d = d[k]
    if value is KeyError:
        del d[path[-1]]
    else:
        d[path[-1]] = value
    return m


@pytest.mark.parametrize("path,value", [
    (("count",), -1), (("count",), 2.5), (("count",), True), (("count",), "4"),
    (("means", "mins"), [0, 0]), (("means", "maxs"), KeyError), (("means", "files"), ["a"]),
    (("scales", "codebook"), KeyError), (("scales", "files"), "c"), (("quats", "files"), ["d", "x"]),
    (("sh0", "codebook"), [0.0, "x"]), (("shN",), {"bands": 1, "codebook": [0.0] * 256}),
    (("shN",), {"bands": -1, "codebook": [], "files": ["f", "g"]}),
])
def test_sog_meta_schema_rejections(path, value):
    """sog.rs:8-42 + 99-101: every departure from the serde schema (missing
    or mistyped fields, a negative, fractional or boolean count, arrays of
    the wrong length) is 'invalid meta.json', before any sheet is read."""
    from splatfield_amd import scene as S
    with pytest.raises(ValueError, match="^invalid meta.json$"):
        S.read_sog(_sog_with_meta(_meta_edit(path, value)))
    # the unedited schema passes the meta stage (it fails later on a sheet)
    with pytest.raises(ValueError, match="missing a"):
        S.read_sog(_sog_with_meta(_SOG_META))
    # null shN is the absent one
    with pytest.raises(ValueError, match="missing a"):
        S.read_sog(_sog_with_meta(dict(_SOG_META, shN=None)))


def test_sog_check_order_and_label_messages():
    """sog.rs:97-198 order: a bad rotation tag is reported before a bad
    bands value (the quats sheet is read first); a label past the palette
    count and a label whose palette leaves the centroid sheet get their own
    messages."""
    import io
    import json
    import zipfile
    from splatfield_amd import scene as S
    from .fixtures import make_sog
    data, _ = make_sog(64, 1, seed=3, bad_tag_at=5)
    z = zipfile.ZipFile(io.BytesIO(data))
    meta = json.loads(z.read("meta.json"))
    meta["shN"]["bands"] = 7
    buf = io.BytesIO()
    with zipfile.ZipFile(buf, "w") as out:
        for name in z.namelist():
            out.writestr(name, json.dumps(meta) if name == "meta.json" else z.read(name))
    with pytest.raises(ValueError, match="quats: invalid rotation tag 17"):
        S.read_sog(buf.getvalue())
    data, _ = make_sog(64, 1, seed=4, bad_label_at=7)
    with pytest.raises(ValueError, match=r"shN label \d+ >= palette size \d+"):
        S.read_sog(data)
    # Off the sheet: with two bands (8 coefficients a palette) make_sog's
    # centroid rows are 16 * 8 + 3 texels wide, 7 rows: 114 palettes by
    # count, 112 on the sheet, so labels 112 and 113 lie past the last row.
    data, sheets = make_sog(64, 2, seed=5)
    info, got, cw, _ = S.read_sog(data)
    cpc, rows = 8, got["centroids"].size // 4 // cw
    count = got["centroids"].size // 4 // cpc
    per_row = cw // cpc
    off = next(l for l in range(count) if l // per_row >= rows)
    z = zipfile.ZipFile(io.BytesIO(data))
    from PIL import Image
    lab = np.asarray(Image.open(io.BytesIO(z.read("shN_labels.webp"))).convert("RGBA")).copy()
    flat = lab.reshape(-1, 4)
    flat[3, 0], flat[3, 1] = off & 255, off >> 8
    img = io.BytesIO()
    Image.fromarray(lab, "RGBA").save(img, "WEBP", lossless=True, exact=True)
    buf = io.BytesIO()
    with zipfile.ZipFile(buf, "w") as out:
        for name in z.namelist():
            out.writestr(name, img.getvalue() if name == "shN_labels.webp" else z.read(name))
    with pytest.raises(ValueError, match=f"shN label {off}: its palette runs off"):
        S.read_sog(buf.getvalue())
