"""Device context and device arrays.

Mirrors python/cudolfinx/context.py: one process-wide context on device 0
(the reference retains the CUDA primary context of device 0,
cpp/cudolfinx/common/CUDA.cpp:28-52).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from ._lib import check, lib

_context = None


class CUDAContext:
    """Owns the mf_context handle (HIP device, hipRTC cache directory)."""

    def __init__(self, device: int = 0, cache_dir: str | None = None):
        h = C.c_void_p()
        cd = cache_dir if cache_dir is not None else os.environ.get("MIFEM_CACHE_DIR")
        check(lib().mf_context_create(device, cd.encode() if cd is not None else None, C.byref(h)))
        self.handle = h
        buf = C.create_string_buffer(64)
        check(lib().mf_context_arch(h, buf, 64))
        self.arch = buf.value.decode()

    def synchronize(self, stream=None):
        check(lib().mf_synchronize(self.handle, stream))

    def __del__(self):
        if getattr(self, "handle", None) and lib is not None:
            try:
                lib().mf_context_destroy(self.handle)
            except Exception:
                pass


def get_cuda_context() -> CUDAContext:
    """Return the process-wide context, creating it on first use."""
    global _context
    if _context is None:
        _context = CUDAContext()
    return _context


class DeviceArray:
    """A float64/int32 array in HIP device memory (owned, freed on delete)."""

    def __init__(self, ctx: CUDAContext, n: int, dtype=np.float64):
        self.ctx = ctx
        self.n = int(n)
        self.dtype = np.dtype(dtype)
        self._ptr = C.c_void_p()
        check(lib().mf_alloc(ctx.handle, max(1, self.n) * self.dtype.itemsize, C.byref(self._ptr)))

    @property
    def ptr(self) -> C.c_void_p:
        return self._ptr

    @property
    def address(self) -> int:
        return int(self._ptr.value or 0)

    @classmethod
    def from_numpy(cls, ctx, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(ctx, a.size, a.dtype)
        d.upload(a)
        return d

    def upload(self, a: np.ndarray, stream=None) -> None:
        a = np.ascontiguousarray(a, dtype=self.dtype)
        if a.size != self.n:
            raise ValueError("size mismatch")
        check(lib().mf_memcpy_htod(self.ctx.handle, self._ptr, a.ctypes.data_as(C.c_void_p),
                                   a.nbytes, stream))
        check(lib().mf_synchronize(self.ctx.handle, stream))

    def to_numpy(self, stream=None) -> np.ndarray:
        out = np.empty(self.n, dtype=self.dtype)
        check(lib().mf_memcpy_dtoh(self.ctx.handle, out.ctypes.data_as(C.c_void_p), self._ptr,
                                   out.nbytes, stream))
        return out

    def __del__(self):
        if getattr(self, "_ptr", None) is not None and self._ptr.value:
            try:
                lib().mf_free(self.ctx.handle, self._ptr)
            except Exception:
                pass
            self._ptr = C.c_void_p()
