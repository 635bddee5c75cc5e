"""Device context and buffers (the L0 runtime of the reference:
crates/sort/src/tensor.rs:29-77 GpuTensor and src/lib.rs:20-27's single
ordered stream), over libsfk's C ABI."""
import ctypes as C

import numpy as np

from ._lib import lib


class SfkError(RuntimeError):
    pass


class Context:
    """One device and one HIP stream; every call enqueues in order."""

    def __init__(self, device=0):
        out = C.c_void_p()
        rc = lib().sfk_ctx_create(device, C.byref(out))
        if rc != 0:
            raise SfkError(f"sfk_ctx_create(device={device}) failed ({rc})")
        self.ptr = out.value

    def check(self, rc, what="sfk call"):
        if rc != 0:
            msg = lib().sfk_last_error(self.ptr)
            raise SfkError(f"{what}: {msg.decode() if msg else ''} (rc={rc})")

    def sync(self):
        self.check(lib().sfk_sync(self.ptr), "sync")

    def device_name(self):
        buf = C.create_string_buffer(256)
        self.check(lib().sfk_device_name(self.ptr, buf, 256), "device_name")
        return buf.value.decode()

    def timer_start(self):
        self.check(lib().sfk_timer_start(self.ptr))

    def timer_stop_ms(self):
        ms = C.c_float()
        self.check(lib().sfk_timer_stop_ms(self.ptr, C.byref(ms)))
        return ms.value

    def prof_enable(self, on=True):
        self.check(lib().sfk_prof_enable(self.ptr, int(on)), "prof_enable")

    def prof_report(self):
        """{name: (launches, total_ms)} of the launches recorded since the
        last report (SYNC)."""
        buf = C.create_string_buffer(1 << 16)
        self.check(lib().sfk_prof_report(self.ptr, buf, 1 << 16), "prof_report")
        out = {}
        for line in buf.value.decode().splitlines():
            name, cnt, ms = line.split()
            out[name] = (int(cnt), float(ms))
        return out

    def close(self):
        if self.ptr:
            lib().sfk_ctx_destroy(self.ptr)
            self.ptr = None


class GpuTensor:
    """A device buffer with a numpy dtype and shape.  ``owner`` keeps alive
    the object whose memory a borrowed view aliases."""

    def __init__(self, ctx, ptr, shape, dtype, owned=True, owner=None):
        self.ctx, self.ptr = ctx, ptr
        self.shape = tuple(shape)
        self.dtype = np.dtype(dtype)
        self.owned = owned
        self.owner = owner

    @property
    def nbytes(self):
        return int(np.prod(self.shape, dtype=np.int64)) * self.dtype.itemsize

    @classmethod
    def empty(cls, ctx, shape, dtype=np.uint32):
        if isinstance(shape, int):
            shape = (shape,)
        t = cls(ctx, None, shape, dtype)
        out = C.c_void_p()
        ctx.check(lib().sfk_alloc(ctx.ptr, max(t.nbytes, 4), C.byref(out)), "alloc")
        t.ptr = out.value
        return t

    @classmethod
    def from_numpy(cls, ctx, arr):
        arr = np.ascontiguousarray(arr)
        t = cls.empty(ctx, arr.shape, arr.dtype)
        if arr.nbytes:
            ctx.check(lib().sfk_upload(ctx.ptr, t.ptr, arr.ctypes.data, arr.nbytes), "upload")
        return t

    def write(self, arr):
        arr = np.ascontiguousarray(arr, dtype=self.dtype)
        assert arr.nbytes <= self.nbytes
        self.ctx.check(lib().sfk_upload(self.ctx.ptr, self.ptr, arr.ctypes.data, arr.nbytes))

    def read(self, count=None):
        n = int(np.prod(self.shape, dtype=np.int64)) if count is None else count
        out = np.empty(n, self.dtype)
        if n:
            self.ctx.check(lib().sfk_download(self.ctx.ptr, out.ctypes.data, self.ptr,
                                              out.nbytes), "download")
        return out if count is not None else out.reshape(self.shape)

    def fill_u32(self, value=0):
        count = self.nbytes // 4
        self.ctx.check(lib().sfk_fill_u32(self.ctx.ptr, self.ptr, value, count), "fill")

    def view(self, ptr, shape, dtype=None):
        return GpuTensor(self.ctx, ptr, shape, dtype or self.dtype, owned=False, owner=self)

    def free(self):
        if self.owned and self.ptr:
            lib().sfk_free(self.ctx.ptr, self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            if self.owned and self.ptr and self.ctx.ptr:
                lib().sfk_free(self.ctx.ptr, self.ptr)
        except Exception:
            pass
