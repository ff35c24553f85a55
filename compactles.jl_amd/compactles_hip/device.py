"""Device context and device arrays on top of the C ABI (no torch needed)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import check, load


class Context:
    """One HIP device and its stream (cles_ctx)."""

    def __init__(self, device: int = 0):
        self.lib = load()
        h = C.c_void_p()
        check(self.lib.cles_ctx_create(device, C.byref(h)))
        self.handle = h
        self.device = device

    def close(self):
        if getattr(self, "handle", None):
            self.lib.cles_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def synchronize(self):
        check(self.lib.cles_synchronize(self.handle))


class DeviceArray:
    """A column-major (Fortran-order) device array of float64 or float32."""

    def __init__(self, ctx: Context, shape, dtype=np.float64):
        self.ctx = ctx
        self.shape = tuple(int(s) for s in shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        check(ctx.lib.cles_malloc(ctx.handle, self.nbytes, C.byref(p)))
        self.ptr = p

    @classmethod
    def from_numpy(cls, ctx: Context, a: np.ndarray, dtype=None):
        dt = np.dtype(dtype) if dtype is not None else a.dtype
        a = np.asfortranarray(a, dtype=dt)
        d = cls(ctx, a.shape, dt)
        d.upload(a)
        return d

    def upload(self, a: np.ndarray):
        a = np.asfortranarray(a, dtype=self.dtype)
        assert a.shape == self.shape
        check(self.ctx.lib.cles_memcpy_h2d(self.ctx.handle, self.ptr, a.ctypes.data_as(C.c_void_p), self.nbytes))

    def numpy(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=self.dtype, order="F")
        check(self.ctx.lib.cles_memcpy_d2h(self.ctx.handle, out.ctypes.data_as(C.c_void_p), self.ptr, self.nbytes))
        return out

    def zero(self):
        check(self.ctx.lib.cles_memset(self.ctx.handle, self.ptr, 0, self.nbytes))

    def free(self):
        if getattr(self, "ptr", None):
            self.ctx.lib.cles_free(self.ctx.handle, self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
