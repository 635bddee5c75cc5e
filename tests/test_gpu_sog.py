"""GPU SOG ingestion (src/sog.rs:97-198): the device decode of the sheets
equals the oracle's numpy decode (codebook lookups, quaternions, palettes
bit for bit; the exp_m1 means and logit opacities to 1 ulp, both sides
rounding an f64 evaluation), and decode failures carry the reference's
messages."""
import numpy as np
import pytest

from oracle import oracle as O
from splatfield_amd import render as R
from splatfield_amd import scene as S
from splatfield_amd.runtime import GpuTensor

from .fixtures import make_sog

pytestmark = pytest.mark.gpu


def ulps(a, b):
    a = np.asarray(a, np.float32).view(np.int32).astype(np.int64)
    b = np.asarray(b, np.float32).view(np.int32).astype(np.int64)
    return np.abs(a - b)


@pytest.mark.parametrize("n,bands", [(1, 0), (1000, 1), (4099, 3)])
def test_from_sog_matches_oracle(ctx, n, bands):
    data, _ = make_sog(n, bands, seed=10 + bands)
    sp = R.Splats.from_sog(ctx, data)
    per = (bands + 1) ** 2
    assert (sp.n, sp.sh_per_ch) == (n, per)
    ra, rs = O.sog_decode(*S.read_sog(data))
    d_attrs, d_sh = sp.device_ptrs()
    ga = GpuTensor(ctx, d_attrs, (11 * n,), np.float32, False, sp).read().reshape(11, n)
    gs = GpuTensor(ctx, d_sh, (3 * per * n,), np.float32, False, sp).read()
    ra = ra.reshape(11, n)
    np.testing.assert_array_equal(ga[3:10], ra[3:10])
    np.testing.assert_array_equal(gs, rs)
    assert ulps(ga[:3], ra[:3]).max() <= 1 and ulps(ga[10], ra[10]).max() <= 1
    assert np.mean(ga[:3] == ra[:3]) > 0.999


def test_from_sog_errors(ctx):
    data, _ = make_sog(300, 1, seed=3, bad_tag_at=123)
    with pytest.raises(ValueError, match="quats: invalid rotation tag 17"):
        R.Splats.from_sog(ctx, data)
    data, _ = make_sog(300, 2, seed=4, bad_label_at=77)
    with pytest.raises(ValueError, match="shN label .* >= palette size"):
        R.Splats.from_sog(ctx, data)


def test_sog_scene_renders(ctx):
    """A decoded SOG scene goes through the render path like any other."""
    data, _ = make_sog(4096, 3, seed=5)
    sp = R.Splats.from_sog(ctx, data)
    cam = R.Camera.look_at((0.0, 0.0, 12.0), (0.0, 0.0, 0.0))
    sc = R.RenderScratch(ctx, sp.n, 128, 128)
    a = R.render_with(sp, sc, cam, 128, 128)
    b = R.render_with(sp, sc, cam, 128, 128)
    np.testing.assert_array_equal(a, b)
