"""The active segmentation loop (src/seg/active.rs:185-232) over the device
pipeline, with the reference's own loop tests (active.rs:292-496): the
three-cluster scene of views.rs:113-131 and its perfect sphere oracle, 256^2
renders, 12 candidate views, the bootstrap camera framing the scene."""
import numpy as np
import pytest

from splatfield_amd import render as R
from splatfield_amd import seg as G

from .fixtures import PER_CLUSTER, cluster_scene, sphere_oracle

pytestmark = pytest.mark.gpu


def harness(ctx, iterations, batched=True, camera=None):
    """active.rs:292-315."""
    attrs, centers = cluster_scene()
    n = 3 * PER_CLUSTER
    splats = R.Splats(ctx, attrs, np.zeros(3 * n, np.float32))
    if camera is None:
        camera = R.Camera()
        camera.frame_bounds(splats.bounds)
    cfg = G.Config(resolution=256, candidates=12, iterations=iterations)
    return G.Segmenter(splats, cfg, camera, batched=batched), centers


@pytest.mark.parametrize("batched", [True, False], ids=["batched", "plain"])
def test_active_loop_segments_target_cluster(ctx, batched):
    """active.rs:376-419: six rounds with a perfect oracle isolate cluster 0:
    every round after the blind one scores a positive EIG, the MAP mask is
    right for >= 90% of the splats, the target's posterior means average
    above 0.7 and no other splat reaches 0.5.  The batched round and the plain
    per-candidate loop both meet it."""
    seg, centers = harness(ctx, 6, batched=batched)
    rounds = seg.run_with(sphere_oracle([centers[0]]), lambda r, s: True)
    assert len(rounds) == 6
    assert all(r.eig > 0.0 for r in rounds[1:]), [r.eig for r in rounds]
    labels = seg.segmentation()
    n = 3 * PER_CLUSTER
    correct = int(np.sum(labels == (np.arange(n) < PER_CLUSTER)))
    assert correct >= n * 90 // 100, correct
    a, b = seg.posteriors()
    m = a / (a + b)
    assert m[:PER_CLUSTER].mean() > 0.7, m[:PER_CLUSTER].mean()
    assert m[PER_CLUSTER:].max() < 0.5, m[PER_CLUSTER:].max()


def test_run_with_stops_when_on_round_returns_false(ctx):
    """active.rs:350-374: on_round returning False after the third round
    stops the run there; the posterior so far has moved off the prior."""
    seg, centers = harness(ctx, 6)
    seen = []

    def on_round(r, state):
        seen.append(r)
        return len(seen) < 3
    rounds = seg.run_with(sphere_oracle([centers[0]]), on_round)
    assert len(rounds) == 3 and len(seen) == 3
    a, b = seg.posteriors()
    assert np.any((np.abs(a - 1) > 1e-3) | (np.abs(b - 1) > 1e-3))


def test_selected_views_face_the_object(ctx):
    """active.rs:421-437: every EIG-selected camera faces the object's
    hemisphere."""
    seg, centers = harness(ctx, 3)
    rounds = seg.run_with(sphere_oracle([centers[0]]))
    assert len(rounds) == 3
    for r in rounds[1:]:
        to_obj = np.asarray(centers[0][0], np.float64) - np.asarray(r.camera.position)
        assert r.camera.forward() @ (to_obj / np.linalg.norm(to_obj)) > 0.0


def test_initial_camera_is_honored(ctx):
    """active.rs:439-462: round 0 observes from the camera handed to the
    constructor, exactly; look_at from 3 behind the target along -Z is the
    identity arc."""
    _, centers = cluster_scene()
    c = centers[0][0]
    cam = R.Camera.look_at(c - np.array([0, 0, 3], np.float32), c)
    assert cam.rotation == (0.0, 0.0, 0.0, 1.0)
    seg, _ = harness(ctx, 6, camera=cam)
    rounds = seg.run_with(sphere_oracle([centers[0]]))
    assert rounds[0].camera.position == cam.position
    assert rounds[0].camera.rotation == cam.rotation
    assert rounds[0].eig == 0.0
    assert len(rounds) == 6


def test_blind_oracle_stops_after_one_round(ctx):
    """active.rs:464-473: an oracle that never sees anything ends the run
    after its first (empty) mask."""
    seg, _ = harness(ctx, 6)
    rounds = seg.run_with(sphere_oracle([]))
    assert len(rounds) == 1 and rounds[0].fg_pixels == 0


def test_unproductive_view_is_not_replayed(ctx):
    """active.rs:475-496: a one-pixel mask fires the oracle but tips no splat
    past a > b, so round 1 surveys the whole scene from a new camera, with a
    positive EIG."""
    def sliver(rgb, size, camera):
        m = np.zeros(size[0] * size[1], np.uint8)
        m[0] = 1
        return m
    seg, _ = harness(ctx, 3)
    rounds = seg.run_with(sliver)
    assert len(rounds) == 3
    assert rounds[0].fg_pixels == 1
    assert rounds[1].camera.position != rounds[0].camera.position
    assert rounds[1].eig > 0.0


def test_oracle_errors_surface(ctx):
    """active.rs:167-168: the oracle's failure ends the step as an error."""
    seg, _ = harness(ctx, 2)

    def broken(rgb, size, camera):
        raise OSError("sensor offline")
    with pytest.raises(RuntimeError, match="oracle failed"):
        seg.run_with(broken)
