"""crates/sort mirror: RadixScratch.argsort, ScanScratch and
exclusive_scan_gather (crates/sort/src/sort.rs, scan.rs) over libsfk."""
import ctypes as C

import numpy as np

from ._lib import lib
from .runtime import GpuTensor


def bits_for(max_exclusive):
    """Bits needed to represent any value in 0..max_exclusive (sort.rs:259-261)."""
    return int(lib().sfk_bits_for(max_exclusive))


class RadixScratch:
    """Reusable ping-pong buffers for argsort (sort.rs:278-300).  Outputs of
    ``argsort`` alias this scratch until its next use."""

    def __init__(self, ctx, max_elems, key_bytes=4, digit_bits=8):
        self.ctx = ctx
        self.key_bytes = key_bytes
        self.max_elems = max_elems
        out = C.c_void_p()
        ctx.check(lib().sfk_radix_scratch_create(ctx.ptr, max_elems, key_bytes, C.byref(out)),
                  "radix scratch")
        self.ptr = out.value
        self.set_digit_bits(digit_bits)

    def set_digit_bits(self, bits):
        self.ctx.check(lib().sfk_radix_set_digit_bits(self.ptr, bits), "digit bits")

    def set_bucket_finish(self, on):
        """u64 sorts of >= 2^20 keys: two passes on the top 16 bits, then each
        group sorted on chip when all fit (default on); same results."""
        self.ctx.check(lib().sfk_radix_set_bucket_finish(self.ptr, 1 if on else 0),
                       "bucket finish")

    def argsort(self, keys, vals, n, bits, write_keys=True, d_n=None):
        """Stable argsort of keys[0..n) on their low ``bits`` (sort.rs:302-381).
        ``vals`` may be None (identity).  Returns (keys, vals) GpuTensors that
        alias the scratch (or the inputs when nothing moves)."""
        ok, ov = C.c_void_p(), C.c_void_p()
        fn = lib().sfk_argsort_u32 if self.key_bytes == 4 else lib().sfk_argsort_u64
        self.ctx.check(fn(self.ctx.ptr, self.ptr, keys.ptr, vals.ptr if vals is not None else None,
                          n, d_n.ptr if d_n is not None else None, bits, int(write_keys),
                          C.byref(ok), C.byref(ov)), "argsort")
        kdt = np.uint32 if self.key_bytes == 4 else np.uint64
        return (GpuTensor(self.ctx, ok.value, (n,), kdt, owned=False, owner=self),
                GpuTensor(self.ctx, ov.value, (n,), np.uint32, owned=False, owner=self))

    def argsort_segmented(self, keys, nseg, seg_len, bits, write_keys=True):
        """nseg independent stable argsorts of seg_len keys each, back to back
        (the batched multi-view depth sort; no reference counterpart).  Values
        are indices within each segment.  Returns (keys, vals) GpuTensors that
        alias the scratch."""
        ok, ov = C.c_void_p(), C.c_void_p()
        self.ctx.check(lib().sfk_argsort_u32_segmented(self.ctx.ptr, self.ptr, keys.ptr, None,
                                                       nseg, seg_len, bits, int(write_keys),
                                                       C.byref(ok), C.byref(ov)),
                       "argsort segmented")
        n = nseg * seg_len
        return (GpuTensor(self.ctx, ok.value, (n,), np.uint32, owned=False, owner=self),
                GpuTensor(self.ctx, ov.value, (n,), np.uint32, owned=False, owner=self))

    def __del__(self):
        try:
            if self.ctx.ptr:  # the context must outlive the objects it made
                lib().sfk_radix_scratch_destroy(self.ptr)
        except Exception:
            pass


class ScanScratch:
    """Per-block totals of the scan (scan.rs:47-60)."""

    def __init__(self, ctx, max_elems):
        self.ctx = ctx
        out = C.c_void_p()
        ctx.check(lib().sfk_scan_scratch_create(ctx.ptr, max_elems, C.byref(out)), "scan scratch")
        self.ptr = out.value

    def __del__(self):
        try:
            if self.ctx.ptr:  # the context must outlive the objects it made
                lib().sfk_scan_scratch_destroy(self.ptr)
        except Exception:
            pass


def exclusive_scan_gather(gids, counts, offsets, n, scratch):
    """offsets[i] = sum_{j<i} counts[gids[j]] (scan.rs:215-255)."""
    ctx = scratch.ctx
    ctx.check(lib().sfk_exclusive_scan_gather(ctx.ptr, gids.ptr, counts.ptr, offsets.ptr, n,
                                              scratch.ptr), "exclusive_scan_gather")


def exclusive_scan_buf(buf, n, scratch):
    """In-place exclusive scan of buf[0..n) (scan.rs:182-213)."""
    ctx = scratch.ctx
    ctx.check(lib().sfk_exclusive_scan(ctx.ptr, buf.ptr, n, scratch.ptr), "exclusive_scan")
