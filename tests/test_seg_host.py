"""Host-side logic of the segmentation loop (CPU): config validation before
any device work, the bitmap unpack, the MAP predicate, the candidate sphere's
closed form and the bootstrap camera framing."""
import math

import numpy as np
import pytest

from splatfield_amd import render as R
from splatfield_amd import seg as G


def test_segmenter_rejects_invalid_config():
    """active.rs:317-348: zero candidates, a resolution below 16 and zero
    iterations are rejected by the constructor before any device work (no
    splats are even needed)."""
    base = G.Config(resolution=256, candidates=12, iterations=2)
    for cfg, msg in [(G.Config(256, 0, 2), "candidates must be >= 1"),
                     (G.Config(8, 12, 2), "resolution must be >= 16"),
                     (G.Config(256, 12, 0), "iterations must be >= 1")]:
        with pytest.raises(ValueError, match=msg):
            G.Segmenter(None, cfg, R.Camera())
    assert base.resolution == 256
    assert G.Config() == G.Config(512, 20, 20)  # active.rs:45-55


def test_bitmap_to_rgb_handles_row_padding():
    """active.rs:262-290: a 2 x 5-word bitmap with 3 valid pixels per row
    yields exactly the 18 RGB bytes, alpha and padding dropped."""
    words = np.array([0x04030201, 0x08070605, 0x0C0B0A09, 0xDEADBEEF, 0xDEADBEEF,
                      0x14131211, 0x18171615, 0x1C1B1A19, 0xDEADBEEF, 0xDEADBEEF], np.uint32)
    rgb = G.bitmap_to_rgb(words, 3, 2, 5)
    want = [b for b in list(range(1, 13)) + list(range(17, 29)) if b % 4 != 0]
    assert rgb.tolist() == want


def test_map_labels_ties_are_background():
    """beta.rs:317-324, and the length check of beta.rs:294-297."""
    assert G.map_labels([1.0, 2.0, 0.5], [1.0, 1.0, 0.5]).tolist() == [False, True, False]
    with pytest.raises(ValueError):
        G.map_labels([1.0, 2.0], [1.0])


def test_candidates_on_sphere_looking_at_center():
    """views.rs:279-311: 20 cameras at distance 1.5 r / tan(fov / 2) from the
    centre (1e-4), each looking at it, Fibonacci z strictly decreasing."""
    center = np.array([1.0, 2.0, 3.0], np.float32)
    radius = 0.5
    want = 1.5 * radius / math.tan(G.CANDIDATE_FOV * 0.5)
    cams = G.candidates(center, radius, 20)
    assert len(cams) == 20
    zs = []
    for k, cam in enumerate(cams):
        off = np.asarray(cam.position, np.float64) - center
        assert abs(np.linalg.norm(off) - want) < 1e-4, k
        to_center = -off / np.linalg.norm(off)
        fwd = cam.forward()
        assert fwd @ to_center / np.linalg.norm(fwd) > 1 - 1e-4, k
        assert tuple(cam.fov) == (G.CANDIDATE_FOV, G.CANDIDATE_FOV)
        zs.append((off / np.linalg.norm(off))[2])
    assert all(a > b for a, b in zip(zs, zs[1:]))


def test_candidates_clamp_degenerate_radius():
    """views.rs:199-226: a one-splat localization (radius ~0) is lifted to
    the 1e-3 floor: every camera at 1.5e-3 / tan(fov / 2), finite."""
    center = np.array([2.0, -1.0, 0.5], np.float32)
    want = 1.5 * 1e-3 / math.tan(G.CANDIDATE_FOV * 0.5)
    for cam in G.candidates(center, 0.0, 5):
        off = np.asarray(cam.position, np.float64) - center
        assert abs(np.linalg.norm(off) - want) < 1e-6
        assert np.all(np.isfinite(cam.forward())) and np.all(np.isfinite(cam.position))


def test_frame_bounds_looks_along_plus_y():
    """camera.rs:46-52: centred on the bounds, pulled back along -Y by twice
    the largest extent, looking along +Y."""
    cam = R.Camera()
    d = cam.frame_bounds((np.array([-1, -2, 0], np.float32), np.array([3, 0, 1], np.float32)))
    assert d == 8.0
    np.testing.assert_allclose(cam.position, (1.0, -9.0, 0.5))
    np.testing.assert_allclose(cam.forward(), (0, 1, 0), atol=1e-6)
