"""Device buffers and copies through the C ABI (CuAllocateArray / CuCopyArray*)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._abi import check, load


class DevBuf:
    """A device allocation made with CuAllocateArray; freed with CuFreeArray."""

    def __init__(self, nbytes: int, lib=None):
        self.lib = lib or load()
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        self.lib.CuAllocateArray(C.byref(p), max(self.nbytes, 4))
        check(self.lib)
        self.ptr = p.value

    @classmethod
    def from_array(cls, a: np.ndarray, lib=None) -> "DevBuf":
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes, lib)
        b.upload(a)
        return b

    def upload(self, a: np.ndarray, offset: int = 0) -> None:
        a = np.ascontiguousarray(a)
        assert offset + a.nbytes <= self.nbytes, "upload past the end of the buffer"
        self.lib.CuCopyArrayToDevice(self.ptr, a.ctypes.data, offset, a.nbytes)
        check(self.lib)

    def download(self, dtype, count: int | None = None) -> np.ndarray:
        dtype = np.dtype(dtype)
        if count is None:
            count = self.nbytes // dtype.itemsize
        out = np.empty(count, dtype=dtype)
        assert out.nbytes <= self.nbytes, "download past the end of the buffer"
        self.lib.CuCopyArrayFromDevice(out.ctypes.data, self.ptr, None, out.nbytes)
        check(self.lib)
        return out

    def fill(self, byte: int) -> None:
        self.lib.CuSetArrayValue(self.ptr, byte, self.nbytes)
        check(self.lib)

    def free(self) -> None:
        if self.ptr:
            self.lib.CuFreeArray(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
