"""Batched Linked List frames (oit_batch_*): many frames of different size,
depth complexity and K in one resident pool, one append launch for all of
them and one resolve launch per (K, tail_blend) group, optionally on forked
streams or replayed from a captured hipGraph, with an autotuner over the
kernel variants. Each frame's image equals the single-frame Linked List
result (pool >= fragments + 1 nodes) bit for bit.

No counterpart in the reference, which draws one frame per command buffer
(oitRender.cpp:28-154); the per-frame work is the Linked List pair of
clearTransparent/drawTransparent (oitRender.cpp:201-227).
"""
from __future__ import annotations

import ctypes
import dataclasses

from . import _abi
from .device import stream_ptr


def _dptr(b) -> int:
    if b is None:
        return 0
    return b if isinstance(b, int) else int(b.ptr)


@dataclasses.dataclass
class BatchFrame:
    """One frame: its size, K, tail flag and device buffers (fragments in
    draw order with frame-local targets; output [width*height*samples, 4])."""

    width: int
    height: int
    layers: int
    target: object
    depth: object
    rgba: object
    count: int
    out: object
    samples: int = 1
    tail_blend: bool = True
    background: tuple[float, float, float, float] = (0.5, 0.5, 0.5, 1.0)

    def to_c(self) -> _abi.OitBatchFrame:
        f = _abi.OitBatchFrame()
        f.width, f.height, f.samples = self.width, self.height, self.samples
        f.layers = self.layers
        f.tail_blend = 1 if self.tail_blend else 0
        f.reserved = 0
        for i in range(4):
            f.background[i] = self.background[i]
        f.frags.target = _dptr(self.target)
        f.frags.depth = _dptr(self.depth)
        f.frags.rgba = _dptr(self.rgba)
        f.frags.count = int(self.count)
        f.out = _dptr(self.out)
        return f


def _frames_c(frames):
    arr = (_abi.OitBatchFrame * max(len(frames), 1))()
    for i, fr in enumerate(frames):
        arr[i] = fr.to_c()
    return arr, len(frames)


class Batch:
    def __init__(self, device: int = 0, concurrent: bool = False, graph: bool = False, frame_order: bool = False):
        self.handle = ctypes.c_void_p(0)
        flags = ((_abi.BATCH_CONCURRENT if concurrent else 0) | (_abi.BATCH_GRAPH if graph else 0)
                 | (_abi.BATCH_FRAME_ORDER if frame_order else 0))
        _abi.check("oit_batch_create", _abi.lib().oit_batch_create(ctypes.byref(self.handle), device, flags))

    def query(self, frames) -> _abi.OitBatchSizes:
        arr, n = _frames_c(frames)
        out = _abi.OitBatchSizes()
        _abi.check("oit_batch_query", _abi.lib().oit_batch_query(self.handle, arr, n, ctypes.byref(out)))
        return out

    def reserve(self, pool_nodes: int, head_entries: int) -> None:
        _abi.check("oit_batch_reserve", _abi.lib().oit_batch_reserve(self.handle, pool_nodes, head_entries))

    def run(self, frames, stream=None) -> None:
        arr, n = _frames_c(frames)
        _abi.check("oit_batch_run", _abi.lib().oit_batch_run(self.handle, arr, n, stream_ptr(stream)))

    def tune(self, frames, stream=None) -> None:
        arr, n = _frames_c(frames)
        _abi.check("oit_batch_tune", _abi.lib().oit_batch_tune(self.handle, arr, n, stream_ptr(stream)))

    @property
    def tuning(self) -> _abi.OitBatchTuning:
        t = _abi.OitBatchTuning()
        _abi.check("oit_batch_get_tuning", _abi.lib().oit_batch_get_tuning(self.handle, ctypes.byref(t)))
        return t

    @tuning.setter
    def tuning(self, t: _abi.OitBatchTuning) -> None:
        _abi.check("oit_batch_set_tuning", _abi.lib().oit_batch_set_tuning(self.handle, ctypes.byref(t)))

    def close(self) -> None:
        if self.handle:
            _abi.lib().oit_batch_destroy(self.handle)
            self.handle = ctypes.c_void_p(0)

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def render_noalloc(frame: BatchFrame, stream=None, device: int = 0) -> None:
    """One Linked List frame through oit_ll_render_noalloc: the image of
    oit_clear + oit_append + oit_resolve with a pool of count + 1 nodes, with
    the heads and nodes in the library's own device array (no allocation).
    Raises OitError (OIT_ERR_CAPACITY) for a frame the array cannot hold."""
    f = frame.to_c()
    _abi.check("oit_ll_render_noalloc", _abi.lib().oit_ll_render_noalloc(device, ctypes.byref(f), stream_ptr(stream)))
