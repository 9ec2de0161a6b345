"""Device memory, streams and events through the library's own C ABI.

The package needs no other HIP binding: buffers are raw device pointers
owned by ``DeviceBuffer`` and filled from / read into numpy arrays. (A
second HIP runtime in the same process, e.g. one bundled with another
framework, would share the ``libamdhip64.so.7`` soname and take over the
library's runtime; keeping the GPU path on this module avoids that.)
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _abi


def device_count() -> int:
    n = ctypes.c_int(0)
    rc = _abi.lib().oit_device_count(ctypes.byref(n))
    return int(n.value) if rc == 0 else 0


class Stream:
    def __init__(self):
        self.handle = ctypes.c_void_p(0)
        _abi.check("oit_stream_create", _abi.lib().oit_stream_create(ctypes.byref(self.handle)))

    @property
    def ptr(self) -> int:
        return int(self.handle.value or 0)

    def synchronize(self) -> None:
        _abi.check("oit_stream_synchronize", _abi.lib().oit_stream_synchronize(self.handle))

    def close(self) -> None:
        if self.handle:
            _abi.lib().oit_stream_destroy(self.handle)
            self.handle = ctypes.c_void_p(0)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Event:
    def __init__(self):
        self.handle = ctypes.c_void_p(0)
        _abi.check("oit_event_create", _abi.lib().oit_event_create(ctypes.byref(self.handle)))

    def record(self, stream: "Stream | None" = None) -> None:
        _abi.check("oit_event_record", _abi.lib().oit_event_record(self.handle, stream_ptr(stream)))

    def synchronize(self) -> None:
        _abi.check("oit_event_synchronize", _abi.lib().oit_event_synchronize(self.handle))

    def elapsed_ms(self, end: "Event") -> float:
        ms = ctypes.c_float(0)
        _abi.check("oit_event_elapsed_ms", _abi.lib().oit_event_elapsed_ms(ctypes.byref(ms), self.handle, end.handle))
        return float(ms.value)

    def __del__(self):
        try:
            if self.handle:
                _abi.lib().oit_event_destroy(self.handle)
        except Exception:
            pass


def stream_ptr(stream) -> int:
    if stream is None:
        return 0
    if isinstance(stream, int):
        return stream
    return stream.ptr


class DeviceBuffer:
    """A device allocation with a dtype and shape, for numpy round trips."""

    def __init__(self, shape, dtype):
        self.shape = tuple(int(x) for x in (shape if isinstance(shape, (tuple, list)) else (shape,)))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape, dtype=np.int64)) * self.dtype.itemsize
        self.handle = ctypes.c_void_p(0)
        _abi.check("oit_device_alloc", _abi.lib().oit_device_alloc(ctypes.byref(self.handle), self.nbytes))

    @classmethod
    def from_numpy(cls, a: np.ndarray, stream=None) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        b = cls(a.shape, a.dtype)
        b.upload(a, stream)
        if stream is not None:
            stream.synchronize()
        return b

    @property
    def ptr(self) -> int:
        return int(self.handle.value or 0)

    def upload(self, a: np.ndarray, stream=None) -> None:
        a = np.ascontiguousarray(a, dtype=self.dtype)
        if a.nbytes != self.nbytes:
            raise ValueError("size mismatch")
        _abi.check("oit_memcpy_h2d", _abi.lib().oit_memcpy_h2d(self.handle, a.ctypes.data, self.nbytes,
                                                              stream_ptr(stream)))

    def numpy(self, stream=None) -> np.ndarray:
        out = np.empty(self.shape, dtype=self.dtype)
        _abi.check("oit_memcpy_d2h", _abi.lib().oit_memcpy_d2h(out.ctypes.data, self.handle, self.nbytes,
                                                              stream_ptr(stream)))
        return out

    def view(self, offset_elems: int, count: int) -> int:
        """Raw pointer to element `offset_elems` (for splitting a stream into batches)."""
        return self.ptr + int(offset_elems) * self.dtype.itemsize * int(np.prod(self.shape[1:], dtype=np.int64))

    def free(self) -> None:
        if self.handle:
            _abi.lib().oit_device_free(self.handle)
            self.handle = ctypes.c_void_p(0)

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
