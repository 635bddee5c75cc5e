"""GPU radix argsort and prefix-scan primitives against numpy / the oracle,
following the reference's own tests (crates/sort/src/sort.rs:383-543,
crates/sort/src/scan.rs:278-344) plus the 16M-key u64 standalone workload."""
import numpy as np
import pytest

from oracle import oracle as O
from splatfield_amd import sort as SO
from splatfield_amd.runtime import GpuTensor

pytestmark = pytest.mark.gpu


def argsort_gpu(ctx, keys, vals, bits, write_keys=True, key_bytes=4, digit_bits=8, n=None,
                scratch=None, d_n=None):
    n = len(keys) if n is None else n
    s = scratch or SO.RadixScratch(ctx, max(n, 1), key_bytes, digit_bits)
    kt = GpuTensor.from_numpy(ctx, keys)
    vt = None if vals is None else GpuTensor.from_numpy(ctx, vals)
    ok, ov = s.argsort(kt, vt, n, bits, write_keys, d_n)
    return (ok.read() if write_keys else None), ov.read(), s


@pytest.mark.parametrize("n", [1000, 5000, 200000])
@pytest.mark.parametrize("digit_bits", [8, 4, 6])
def test_sorting_stable(ctx, n, digit_bits):
    """sort.rs:433-453."""
    rng = np.random.default_rng(n + digit_bits)
    keys = rng.integers(0, 37, n).astype(np.uint32)
    ok, ov, _ = argsort_gpu(ctx, keys, np.arange(n, dtype=np.uint32), 8, digit_bits=digit_bits)
    np.testing.assert_array_equal(ov, np.argsort(keys, kind="stable"))
    np.testing.assert_array_equal(ok, np.sort(keys))


@pytest.mark.parametrize("nseg,seg_len,digit_bits", [
    (3, 100003, 8),      # one workgroup per block, ragged last block
    (8, 1 << 20, 8),     # 2048 blocks: two blocks per upsweep workgroup
    (5, 1 << 20, 4),     # 1280 blocks, 4-bit digits
    (9, 700001, 8),      # odd block count: the last workgroup takes one block
])
def test_sorting_segmented(ctx, nseg, seg_len, digit_bits):
    """The batched depth sort's segmented argsort: each segment stably sorted
    on 24-bit keys with ties, values the indices within the segment, against
    numpy's stable argsort per segment."""
    rng = np.random.default_rng(seg_len + nseg)
    keys = rng.integers(0, 1 << 24, nseg * seg_len, dtype=np.uint32)
    keys[rng.random(keys.size) < 0.1] = 12345  # ties
    s = SO.RadixScratch(ctx, nseg * seg_len, 4, digit_bits)
    ok, ov = s.argsort_segmented(GpuTensor.from_numpy(ctx, keys), nseg, seg_len, 24)
    ok, ov = ok.read(), ov.read()
    for g in range(nseg):
        seg = keys[g * seg_len:(g + 1) * seg_len]
        np.testing.assert_array_equal(ov[g * seg_len:(g + 1) * seg_len],
                                      np.argsort(seg, kind="stable"), f"segment {g}")
        np.testing.assert_array_equal(ok[g * seg_len:(g + 1) * seg_len], np.sort(seg))


def test_sort_values_only(ctx):
    """sort.rs:455-473: write_keys = false still yields the stable permutation;
    NULL values give the identity argsort."""
    rng = np.random.default_rng(2)
    keys = rng.integers(0, 2**32, 5000, dtype=np.uint64).astype(np.uint32)
    _, ov, _ = argsort_gpu(ctx, keys, np.arange(5000, dtype=np.uint32), 32, write_keys=False)
    np.testing.assert_array_equal(ov, np.argsort(keys, kind="stable"))
    _, ov, _ = argsort_gpu(ctx, keys, None, 32, write_keys=False)
    np.testing.assert_array_equal(ov, np.argsort(keys, kind="stable"))


@pytest.mark.parametrize("maxv", [2, 64, 1000, 65536])
def test_sorting_partial_bits(ctx, maxv):
    """sort.rs:475-485."""
    rng = np.random.default_rng(3)
    keys = rng.integers(0, maxv, 5000).astype(np.uint32)
    bits = SO.bits_for(maxv)
    ok, ov, _ = argsort_gpu(ctx, keys, np.arange(5000, dtype=np.uint32), bits)
    np.testing.assert_array_equal(ov, np.argsort(keys, kind="stable"))


def test_sorting_32bit_clustered_and_spread(ctx):
    """sort.rs:487-510."""
    rng = np.random.default_rng(4)
    clustered = []
    for i in range(10000):
        s = rng.integers(i, i + 150)
        e = rng.integers(s, s + 250)
        seg = np.arange(s, e)
        clustered.append(seg[rng.random(seg.size) < 0.5])
    clustered = np.concatenate(clustered).astype(np.uint32)
    spread = rng.integers(0, 1_000_000, 500_000).astype(np.uint32)
    for keys in (clustered, spread):
        ok, ov, _ = argsort_gpu(ctx, keys, np.arange(keys.size, dtype=np.uint32), 32)
        np.testing.assert_array_equal(ov, np.argsort(keys, kind="stable"))


def test_sort_subset_of_capacity(ctx):
    """sort.rs:512-542: sort a prefix n of a larger tensor."""
    cap, n = 1_000_000, 1000
    keys = (np.arange(cap, dtype=np.uint64) * 2654435761 % 2**32).astype(np.uint32)
    ok, ov, _ = argsort_gpu(ctx, keys, np.arange(cap, dtype=np.uint32), 32, n=n)
    np.testing.assert_array_equal(ok[:n], np.sort(keys[:n]))
    np.testing.assert_array_equal(keys[ov[:n]], ok[:n])


def test_device_count(ctx):
    """The element count read on the device (min(*d_n, n)) sorts exactly the
    live prefix, with no host sync."""
    rng = np.random.default_rng(5)
    keys = rng.integers(0, 8160, 300000).astype(np.uint32)
    for live in [0, 1, 4097, 123456, 300000, 999999]:
        d_n = GpuTensor.from_numpy(ctx, np.array([live], np.uint32))
        ok, ov, _ = argsort_gpu(ctx, keys, np.arange(keys.size, dtype=np.uint32), 13, d_n=d_n)
        m = min(live, keys.size)
        np.testing.assert_array_equal(ov[:m], np.argsort(keys[:m], kind="stable"))


@pytest.mark.parametrize("n", [1, 2, 4097, 100003])
def test_u64_stable(ctx, n):
    rng = np.random.default_rng(n)
    keys = rng.integers(0, 2**63, n, dtype=np.uint64) * np.uint64(2)
    keys[::3] = keys[0]
    ok, ov, _ = argsort_gpu(ctx, keys, np.arange(n, dtype=np.uint32), 64, key_bytes=8)
    np.testing.assert_array_equal(ov, np.argsort(keys, kind="stable"))
    np.testing.assert_array_equal(ok, np.sort(keys))


def test_u64_16m_standalone(ctx):
    """The standalone workload: 16,777,216 uniform-random u64 keys with u32
    index values: sorted, a permutation, and stable (checked against the
    oracle's counting sort on a 1M-key prefix run)."""
    n = 1 << 24
    rng = np.random.default_rng(0)
    keys = rng.integers(0, 2**64 - 1, n, dtype=np.uint64)
    ok, ov, _ = argsort_gpu(ctx, keys, None, 64, key_bytes=8)
    assert np.all(ok[1:] >= ok[:-1])
    np.testing.assert_array_equal(keys[ov], ok)
    assert np.bincount(ov, minlength=n).max() == 1
    sub = keys[:1 << 20]
    ok2, ov2, _ = argsort_gpu(ctx, sub, None, 64, key_bytes=8)
    _, ref = O.argsort_u64(sub, np.arange(sub.size, dtype=np.uint32), 64)
    np.testing.assert_array_equal(ov2, ref)


def _bucket_keys(kind, n, rng):
    if kind == "uniform64":
        return rng.integers(0, 2**64 - 1, n, dtype=np.uint64), 64
    if kind == "ties":  # 64 keys a bucket, 4 values each: stability inside a bucket
        hi = rng.integers(0, 1 << 16, n, dtype=np.uint64) << np.uint64(48)
        return hi | rng.integers(0, 4, n, dtype=np.uint64), 64
    if kind == "bits40":  # keys < 2^40 sorted on 40 bits: groups on bits 24..39
        return rng.integers(0, 1 << 40, n, dtype=np.uint64), 40
    if kind == "bits24":  # one 8-bit pass inside each group
        return rng.integers(0, 1 << 24, n, dtype=np.uint64), 24
    if kind == "narrow":  # keys < 2^40 sorted on 64 bits: one group -> the plain passes
        return rng.integers(0, 1 << 40, n, dtype=np.uint64), 64
    if kind == "one_heavy":  # 1% of keys in one group: past the on-chip cap -> plain
        k = rng.integers(0, 2**64 - 1, n, dtype=np.uint64)
        k[rng.random(n) < 0.01] = np.uint64(0x1234) << np.uint64(48) | np.uint64(7)
        return k, 64
    raise ValueError(kind)


@pytest.mark.parametrize("kind", ["uniform64", "ties", "bits40", "bits24", "narrow", "one_heavy"])
def test_u64_bucket_finish_matches_stable_argsort(ctx, kind):
    """Two passes on the top 16 bits then an on-chip sort of every group
    (or, when a group exceeds the on-chip capacity, the plain passes): the
    stable argsort of the keys' low `bits`, keys written or not, identity or
    given values; the same as with the strategy switched off."""
    n = 1 << 21
    rng = np.random.default_rng(len(kind))
    keys, bits = _bucket_keys(kind, n, rng)
    ref = np.argsort(keys, kind="stable")
    vals = rng.integers(0, 2**32 - 1, n, dtype=np.uint64).astype(np.uint32)
    s = SO.RadixScratch(ctx, n, 8)
    ok, ov, _ = argsort_gpu(ctx, keys, vals, bits, True, key_bytes=8, scratch=s)
    np.testing.assert_array_equal(ov, vals[ref])
    np.testing.assert_array_equal(ok, keys[ref])
    _, ov, _ = argsort_gpu(ctx, keys, None, bits, False, key_bytes=8, scratch=s)
    np.testing.assert_array_equal(ov, ref)
    s.set_bucket_finish(False)
    _, ov_plain, _ = argsort_gpu(ctx, keys, None, bits, False, key_bytes=8, scratch=s)
    np.testing.assert_array_equal(ov_plain, ref)


@pytest.mark.parametrize("n", [1, 5, 1024, 1025, 70000, 100000, 2_000_003])
def test_scan_gather_matches_cpu(ctx, n):
    """scan.rs:283-311 (gather) and 313-331 (deterministic counts = i % 4)."""
    rng = np.random.default_rng(n)
    counts = (np.arange(n) % 4).astype(np.uint32) if n in (1025, 70000) else \
        rng.integers(0, 20, n).astype(np.uint32)
    gids = rng.permutation(n).astype(np.uint32)
    g = GpuTensor.from_numpy(ctx, gids)
    c = GpuTensor.from_numpy(ctx, counts)
    o = GpuTensor.empty(ctx, n)
    sc = SO.ScanScratch(ctx, n)
    SO.exclusive_scan_gather(g, c, o, n, sc)
    np.testing.assert_array_equal(o.read(), O.scan_gather(gids, counts))
    b = GpuTensor.from_numpy(ctx, counts)
    SO.exclusive_scan_buf(b, n, sc)
    np.testing.assert_array_equal(b.read(), np.concatenate([[0], np.cumsum(counts)[:-1]]))


def test_scan_zero_leaves_output(ctx):
    """scan.rs:333-343."""
    g = GpuTensor.from_numpy(ctx, np.array([0], np.uint32))
    c = GpuTensor.from_numpy(ctx, np.array([7], np.uint32))
    o = GpuTensor.from_numpy(ctx, np.array([0xDEADBEEF], np.uint32))
    SO.exclusive_scan_gather(g, c, o, 0, SO.ScanScratch(ctx, 1))
    assert o.read()[0] == 0xDEADBEEF


def test_undersized_scratch_is_an_error(ctx):
    s = SO.RadixScratch(ctx, 10)
    k = GpuTensor.from_numpy(ctx, np.arange(100, dtype=np.uint32))
    with pytest.raises(Exception, match="undersized"):
        s.argsort(k, None, 100, 8)
