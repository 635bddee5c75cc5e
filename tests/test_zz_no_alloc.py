"""sfk_eig_device_noalloc (fixed device workspace, no allocation) against
sfk_eig_device on the same inputs, through the C ABI: two calls in a row,
the second on fewer splats than the first, bit-equal f64 results (the same
kernels and reduction order), and within 1e-3 nats of the oracle's f64 sum."""
import numpy as np
import pytest

from oracle import oracle as O
from splatfield_amd._lib import lib
from splatfield_amd.runtime import GpuTensor

pytestmark = pytest.mark.gpu

SFK_E_CAPACITY = -2
MAX_SPLATS = 1 << 26


def _posterior(n, seed):
    rng = np.random.default_rng(seed)
    a = (1 + rng.random(n) * 20).astype(np.float32)
    b = (1 + rng.random(n) * 20).astype(np.float32)
    resp = rng.integers(1, 1 << 14, n).astype(np.uint32)
    resp[rng.random(n) < 0.4] = 0
    return a, b, resp


def _eig(ctx, name, a, b, resp, n, scale):
    out = GpuTensor.empty(ctx, 1, np.float64)
    rc = getattr(lib(), name)(ctx.ptr, a.ptr, b.ptr, resp.ptr, n, scale, out.ptr)
    ctx.check(rc, name)
    return float(out.read()[0])


def test_noalloc_eig_matches_eig_device(ctx):
    scale = float(1 << 14)
    for n, seed in ((1_000_003, 1), (70_001, 2)):
        a, b, resp = _posterior(n, seed)
        ad, bd, rd = (GpuTensor.from_numpy(ctx, x) for x in (a, b, resp))
        fixed = _eig(ctx, "sfk_eig_device_noalloc", ad, bd, rd, n, scale)
        grown = _eig(ctx, "sfk_eig_device", ad, bd, rd, n, scale)
        assert fixed == grown, (n, fixed, grown)
        ref = O.eig(a, b, resp, scale)
        assert abs(fixed - ref) <= 1e-3, (n, fixed, ref)
        assert fixed > 0


def test_noalloc_eig_empty_and_oversize(ctx):
    a, b, resp = _posterior(16, 3)
    ad, bd, rd = (GpuTensor.from_numpy(ctx, x) for x in (a, b, resp))
    assert _eig(ctx, "sfk_eig_device_noalloc", ad, bd, rd, 0, 1.0) == 0.0
    out = GpuTensor.empty(ctx, 1, np.float64)
    # rejected on the host before any launch
    rc = lib().sfk_eig_device_noalloc(ctx.ptr, ad.ptr, bd.ptr, rd.ptr, MAX_SPLATS + 1, 1.0,
                                      out.ptr)
    assert rc == SFK_E_CAPACITY
