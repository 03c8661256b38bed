"""Device (HBM) arrays owned through libpvl_hip.so's allocator.

A :class:`DeviceArray` is a typed, C-contiguous block of device memory with a shape.  It
is the resident form of the arrays the reference passes as numpy arrays; host round trips
happen only through :meth:`DeviceArray.from_numpy` and :meth:`DeviceArray.numpy`.
"""

from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib


class DeviceArray:
    """C-contiguous device buffer with a numpy dtype and shape."""

    __slots__ = ("ptr", "shape", "dtype", "nbytes", "_owner")

    def __init__(self, shape, dtype=np.float64, *, _ptr: int | None = None, _owner=None):
        self.shape = tuple(int(s) for s in (shape if isinstance(shape, (tuple, list)) else (shape,)))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape, dtype=np.int64)) * self.dtype.itemsize
        self._owner = _owner
        if _ptr is not None:
            self.ptr = int(_ptr)
            return
        p = C.c_void_p()
        _lib.call("pvl_malloc", C.byref(p), C.c_size_t(self.nbytes))
        self.ptr = int(p.value)

    # -- construction -------------------------------------------------------------------
    @classmethod
    def empty(cls, shape, dtype=np.float64) -> "DeviceArray":
        return cls(shape, dtype)

    @classmethod
    def zeros(cls, shape, dtype=np.float64) -> "DeviceArray":
        out = cls(shape, dtype)
        _lib.call("pvl_memset_zero", C.c_void_p(out.ptr), C.c_size_t(out.nbytes))
        return out

    @classmethod
    def from_numpy(cls, arr) -> "DeviceArray":
        a = np.ascontiguousarray(arr)
        out = cls(a.shape, a.dtype)
        if a.nbytes:
            _lib.call("pvl_memcpy_h2d", C.c_void_p(out.ptr), a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes))
        return out

    # -- access ---------------------------------------------------------------------------
    @property
    def size(self) -> int:
        return int(np.prod(self.shape, dtype=np.int64))

    @property
    def ndim(self) -> int:
        return len(self.shape)

    def numpy(self, out: np.ndarray | None = None) -> np.ndarray:
        """Copy to host (synchronises the library stream)."""
        if out is None:
            out = np.empty(self.shape, self.dtype)
        if self.nbytes:
            _lib.call("pvl_memcpy_d2h", out.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr), C.c_size_t(self.nbytes))
        return out

    def copy_from(self, arr) -> None:
        a = np.ascontiguousarray(arr, dtype=self.dtype)
        if a.nbytes != self.nbytes:
            raise ValueError(f"size mismatch: {a.nbytes} bytes into a {self.nbytes}-byte device array")
        if a.nbytes:
            _lib.call("pvl_memcpy_h2d", C.c_void_p(self.ptr), a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes))

    def view(self, offset_elems: int, shape) -> "DeviceArray":
        """Sub-array sharing memory (no copy)."""
        return DeviceArray(shape, self.dtype, _ptr=self.ptr + offset_elems * self.dtype.itemsize, _owner=self)

    def reshape(self, shape) -> "DeviceArray":
        out = DeviceArray(shape, self.dtype, _ptr=self.ptr, _owner=self)
        if out.nbytes != self.nbytes:
            raise ValueError("reshape changes the size")
        return out

    @property
    def cptr(self) -> C.c_void_p:
        return C.c_void_p(self.ptr)

    def free(self) -> None:
        if self._owner is None and self.ptr:
            _lib.call("pvl_free", C.c_void_p(self.ptr))
        self.ptr = 0

    def __del__(self):
        try:
            if self._owner is None and self.ptr and _lib.alive():
                _lib._LIB.pvl_free(C.c_void_p(self.ptr))
                self.ptr = 0
        except Exception:
            pass

    def __repr__(self) -> str:
        return f"DeviceArray(shape={self.shape}, dtype={self.dtype})"


def synchronize() -> None:
    _lib.call("pvl_synchronize")
