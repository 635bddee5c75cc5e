"""GPU parity of the GUI paint ops (src/seg/paint.rs) against the oracle:
the posterior-mean render meets the render contract (1/255) against the
oracle's paint + RGB raster,
the tint writes exactly the oracle's DC planes, and a save / tint / restore
round trip gives back the original frame word for word."""
import numpy as np
import pytest

from oracle import oracle as O
from splatfield_amd import render as R
from splatfield_amd import scene as S
from splatfield_amd.runtime import GpuTensor

from .fixtures import assert_image_close

pytestmark = pytest.mark.gpu

W = H = 256


def c1_camera():
    fov = np.float32(np.pi / 3)
    return R.Camera.look_at((0.0, 0.0, 3.0), (0.0, 0.0, 0.0), fov=(fov, fov))


def posterior(n, seed):
    rng = np.random.default_rng(seed)
    a = (1.0 + rng.gamma(1.0, 2.0, n)).astype(np.float32)
    b = (1.0 + rng.gamma(1.0, 2.0, n)).astype(np.float32)
    b[::7] = a[::7]  # ties: background for the tint
    return a, b


@pytest.mark.parametrize("per_deg", [0, 3])
def test_render_posterior_matches_oracle(ctx, per_deg):
    """render_posterior (paint.rs:22-47): projected colours replaced by
    a/(a+b) before blending; the image, T and the RGBA8 bitmap within the
    render contract of the oracle's."""
    n = 16384
    attrs, sh, k = S.synthetic_scene(n, per_deg, seed=31 + per_deg)
    a, b = posterior(n, 5)
    cam = c1_camera()
    splats = R.Splats(ctx, attrs, sh)
    scratch = R.RenderScratch(ctx, n, W, H)
    da, db = GpuTensor.from_numpy(ctx, a), GpuTensor.from_numpy(ctx, b)
    bm = splats.render_posterior(scratch, da, db, cam, W, H)
    ref = O.render_posterior(O.View.from_buffer_copy(bytes(cam.view(W, H))), attrs, sh, n, k, W,
                             H, a, b)
    got = scratch.read("rgba")
    assert_image_close(got, scratch.read("final_t"), bm.reshape(-1), ref,
                       scratch.read("n_contrib"))
    px = got.reshape(-1, 4)
    assert np.all(px[:, 0] == px[:, 1]) and np.all(px[:, 1] == px[:, 2]), "grayscale"
    assert px[:, 3].max() > 0.5


def test_render_posterior_wrong_length_renders_colours(ctx):
    """paint.rs:30-32: a posterior sized for another splat count gives the
    normal colour render."""
    n = 4096
    attrs, sh, k = S.synthetic_scene(n, 0, seed=41)
    cam = c1_camera()
    splats = R.Splats(ctx, attrs, sh)
    scratch = R.RenderScratch(ctx, n, W, H)
    a, b = posterior(n - 1, 6)
    bm = splats.render_posterior(scratch, GpuTensor.from_numpy(ctx, a),
                                 GpuTensor.from_numpy(ctx, b), cam, W, H)
    plain = R.render_with(splats, scratch, cam, W, H)
    np.testing.assert_array_equal(bm, plain)


def test_tint_save_restore_round_trip(ctx):
    """tint (paint.rs:49-65, 113-122) writes the oracle's DC planes and
    leaves every other SH plane alone; save_colors / restore_colors
    (paint.rs:67-93) bring back the original frame word for word."""
    n = 8192
    attrs, sh, k = S.synthetic_scene(n, 3, seed=43)
    a, b = posterior(n, 7)
    cam = c1_camera()
    splats = R.Splats(ctx, attrs, sh)
    scratch = R.RenderScratch(ctx, n, W, H)
    before = R.render_with(splats, scratch, cam, W, H)
    saved = splats.save_colors()
    np.testing.assert_array_equal(saved.read().reshape(-1), sh[:3 * n])
    da, db = GpuTensor.from_numpy(ctx, a), GpuTensor.from_numpy(ctx, b)
    color = (0.9, -0.25, 0.125)
    splats.tint(da, db, color)
    _, d_sh = splats.device_ptrs()
    got = GpuTensor(ctx, d_sh, (sh.size,), np.float32, False, splats).read()
    np.testing.assert_array_equal(got, O.tint(sh, a, b, color))
    tinted = R.render_with(splats, scratch, cam, W, H)
    assert np.any(tinted != before)
    splats.restore_colors(saved)
    got = GpuTensor(ctx, d_sh, (sh.size,), np.float32, False, splats).read()
    np.testing.assert_array_equal(got, sh)
    np.testing.assert_array_equal(R.render_with(splats, scratch, cam, W, H), before)
