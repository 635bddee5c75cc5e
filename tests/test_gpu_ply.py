"""GPU scene ingestion (src/ply.rs:9-112): the device decode of the vertex
rows gives the oracle's planes bit for bit, and splats loaded from a PLY
render the same frame as splats built from the oracle's planes."""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O
from splatfield_amd import render as R
from splatfield_amd import scene as S
from splatfield_amd._lib import lib
from splatfield_amd.runtime import GpuTensor

from .test_oracle import _synthetic_ply

pytestmark = pytest.mark.gpu


def test_ply_decode_layout_kat(ctx):
    """ply.rs test_parse_ply_layout's file (one f_rest per channel, listed
    out of order) through sfk_ply_decode."""
    data, rows = _synthetic_ply(3, ["f_rest_0", "f_rest_2", "f_rest_1"])
    lay = R.parse_ply_header(data)
    d_rows = GpuTensor.from_numpy(ctx, rows)
    d_attrs = GpuTensor.empty(ctx, 11 * 3, np.float32)
    d_sh = GpuTensor.empty(ctx, lay.sh_planes * 3, np.float32)
    ctx.check(lib().sfk_ply_decode(ctx.ptr, C.byref(lay), d_rows.ptr, d_attrs.ptr, d_sh.ptr),
              "ply_decode")
    attrs, sh = O.parse_ply(data)
    np.testing.assert_array_equal(d_attrs.read(), attrs)
    np.testing.assert_array_equal(d_sh.read(), sh)


@pytest.mark.parametrize("deg,n", [(0, 1000), (3, 70001)])
def test_splats_from_ply_match_oracle(ctx, deg, n):
    """A written scene (random, unnormalised rotations; extra normal columns)
    loads with every plane equal to the oracle's parse, ragged last block
    included, and renders the oracle-plane scene's frame word for word."""
    attrs, sh, k = S.synthetic_scene(n, deg, seed=50 + deg)
    rng = np.random.default_rng(deg)
    attrs[3 * n:7 * n] *= rng.uniform(0.5, 2.0, 4 * n).astype(np.float32)
    data = O.write_ply(attrs, sh, n)
    sp = R.Splats.from_ply(ctx, data)
    assert (sp.n, sp.sh_per_ch) == (n, k)
    ra, rs = O.parse_ply(data)
    d_attrs, d_sh = sp.device_ptrs()
    np.testing.assert_array_equal(GpuTensor(ctx, d_attrs, (11 * n,), np.float32, False,
                                            sp).read(), ra)
    np.testing.assert_array_equal(GpuTensor(ctx, d_sh, (3 * k * n,), np.float32, False,
                                            sp).read(), rs)
    fov = np.float32(np.pi / 3)
    cam = R.Camera.look_at((0.0, 0.0, 3.0), (0.0, 0.0, 0.0), fov=(fov, fov))
    sc = R.RenderScratch(ctx, n, 128, 128)
    got = R.render_with(sp, sc, cam, 128, 128)
    ref = R.render_with(R.Splats(ctx, ra, rs), sc, cam, 128, 128)
    np.testing.assert_array_equal(got, ref)
    assert (got >> 24).max() > 0


def test_splats_from_ply_errors(ctx):
    """Header errors surface with the reference's message; an SH degree the
    renderer does not take is refused."""
    data, _ = _synthetic_ply(3, [])
    with pytest.raises(Exception, match="Missing property: f_dc_2"):
        R.Splats.from_ply(ctx, data.replace(b"property float f_dc_2\n", b""))
    data, _ = _synthetic_ply(3, ["f_rest_0", "f_rest_1", "f_rest_2"])
    with pytest.raises(Exception, match="SH degree 0-3"):
        R.Splats.from_ply(ctx, data)


def test_ply_decode_from_file_image_at_any_offset(ctx):
    """sfk_ply_decode over a device copy of the whole file (rows at the
    header's data_offset): rows 4-byte but not 16-byte aligned take the
    scalar staging and give the oracle's planes; rows at a byte offset that
    breaks float alignment are refused (SFK_E_ARG), not read."""
    n, deg = 1000, 3
    attrs, sh, k = S.synthetic_scene(n, deg, seed=61)
    data = O.write_ply(attrs, sh, n)
    lay = R.parse_ply_header(data)
    ra, rs = O.parse_ply(data)
    d_attrs = GpuTensor.empty(ctx, 11 * n, np.float32)
    d_sh = GpuTensor.empty(ctx, lay.sh_planes * n, np.float32)
    buf = np.frombuffer(data, np.uint8)
    for mis in (4, 8, 12, 0):
        # place the file so that its rows sit `mis` bytes past a 16-byte boundary
        pad = (mis - lay.data_offset) % 16
        img = np.zeros(pad + buf.size + 16, np.uint8)
        img[pad:pad + buf.size] = buf
        dev = GpuTensor.from_numpy(ctx, img)
        rows = dev.ptr + pad + lay.data_offset
        assert rows % 16 == mis
        ctx.check(lib().sfk_ply_decode(ctx.ptr, C.byref(lay), rows, d_attrs.ptr, d_sh.ptr),
                  "ply_decode")
        np.testing.assert_array_equal(d_attrs.read(), ra, f"rows at +{mis}")
        np.testing.assert_array_equal(d_sh.read(), rs, f"rows at +{mis}")
    for mis in (1, 2, 3):
        assert lib().sfk_ply_decode(ctx.ptr, C.byref(lay), dev.ptr + 16 + mis, d_attrs.ptr,
                                    d_sh.ptr) == -1
