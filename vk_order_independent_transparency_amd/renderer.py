"""Host-side mirror of the reference's renderer interface, over the C ABI.

``State`` carries the reference's settings under its own names (oit.h:64-116)
and derives the antialiasing fields exactly as
``State::recomputeAntialiasingSettings`` does (oit.h:84-115). ``Renderer``
is the per-frame surface of ``Sample``: ``clear`` (clearTransparent<Alg>),
``append`` (the COLOR draw of drawTransparent<Alg>), ``resolve`` (the
COMPOSITE draw), dispatched on ``State.algorithm`` inside the native library.

Arguments are device buffers (``device.DeviceBuffer``) or raw device
pointers, and streams are ``device.Stream`` objects or None (the null
stream); all of it is allocated through the library's own C ABI.
"""
from __future__ import annotations

import ctypes
import dataclasses

from . import _abi
from .device import stream_ptr

ALGORITHM_NAMES = ("simple", "linkedlist", "loop", "loop64", "spinlock", "interlock", "weighted")


@dataclasses.dataclass
class State:
    """Renderer settings; field names and defaults follow oit.h:64-76."""

    algorithm: int = _abi.OIT_SPINLOCK
    oitLayers: int = 8
    linkedListAllocatedPerElement: int = 10
    tailBlend: bool = True
    aaType: int = _abi.AA_NONE
    # implicit, recomputed from aaType (oit.h:78-82)
    msaa: int = 1
    sampleShading: bool = False
    supersample: int = 1

    def coverageShading(self) -> bool:  # oit.h:82
        return self.msaa > 1 and not self.sampleShading

    def recomputeAntialiasingSettings(self) -> None:  # oit.h:84-115
        self.sampleShading = False
        self.supersample = 1
        if self.aaType == _abi.AA_NONE:
            self.msaa = 1
        elif self.aaType == _abi.AA_MSAA_4X:
            self.msaa = 4
        elif self.aaType == _abi.AA_SSAA_4X:
            self.msaa = 4
            self.sampleShading = True
        elif self.aaType == _abi.AA_SUPER_4X:
            self.msaa = 1
            self.supersample = 2
        elif self.aaType == _abi.AA_MSAA_8X:
            self.msaa = 8
        elif self.aaType == _abi.AA_SSAA_8X:
            self.msaa = 8
            self.sampleShading = True
        else:
            raise ValueError(f"antialiasing mode {self.aaType} not implemented")


@dataclasses.dataclass
class FrameConfig:
    """Everything oit_configure needs, from a State plus the viewport."""

    algorithm: int
    layers: int
    tail_blend: bool
    width: int
    height: int
    samples: int = 1
    ll_nodes_per_pixel: int = 10
    ll_pool_nodes: int = 0
    max_fragments: int = 0
    background: tuple[float, float, float, float] = (0.5, 0.5, 0.5, 1.0)
    wboit_view_near: float = 2.0
    wboit_view_span: float = 8.0
    variant: int = 0
    coverage_samples: int = 0

    @classmethod
    def from_state(cls, state: State, width: int, height: int, **kw) -> "FrameConfig":
        state.recomputeAntialiasingSettings()
        w, h = width * state.supersample, height * state.supersample
        cov = state.msaa if state.coverageShading() else 0  # OIT_COVERAGE_SHADING, oit.h:82
        return cls(
            algorithm=state.algorithm,
            layers=state.oitLayers,
            tail_blend=state.tailBlend,
            width=w,
            height=h,
            samples=state.msaa if state.sampleShading else 1,
            ll_nodes_per_pixel=state.linkedListAllocatedPerElement,
            coverage_samples=cov,
            **kw,
        )

    @property
    def targets(self) -> int:
        return self.width * self.height * self.samples

    def to_c(self) -> _abi.OitConfig:
        c = _abi.OitConfig()
        c.algorithm = self.algorithm
        c.layers = self.layers
        c.tail_blend = 1 if self.tail_blend else 0
        c.width = self.width
        c.height = self.height
        c.samples = self.samples
        c.ll_nodes_per_pixel = self.ll_nodes_per_pixel
        c.ll_pool_nodes = self.ll_pool_nodes
        c.max_fragments = self.max_fragments
        for i in range(4):
            c.background[i] = self.background[i]
        c.wboit_view_near = self.wboit_view_near
        c.wboit_view_span = self.wboit_view_span
        c.variant = self.variant
        c.coverage_samples = self.coverage_samples
        c.reserved = 0
        return c


def _stream_ptr(stream) -> int:
    return stream_ptr(stream)


def _dptr(t) -> int:
    """Device address of a DeviceBuffer (or anything with .ptr) or a raw int."""
    if isinstance(t, int):
        return t
    p = getattr(t, "ptr", None)
    if p is None:
        raise TypeError("expected a DeviceBuffer or a device address")
    return int(p)


def _rows(t) -> int:
    return int(t.shape[0])


class Renderer:
    """One native OIT context (oit_ctx) on one device."""

    def __init__(self, device: int = 0):
        self._lib = _abi.lib()
        self._ctx = ctypes.c_void_p(0)
        _abi.check("oit_create", self._lib.oit_create(ctypes.byref(self._ctx), int(device)))
        self.device = device
        self.config: FrameConfig | None = None

    def close(self) -> None:
        if self._ctx:
            self._lib.oit_destroy(self._ctx)
            self._ctx = ctypes.c_void_p(0)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def configure(self, cfg: FrameConfig) -> None:
        c = cfg.to_c()
        _abi.check("oit_configure", self._lib.oit_configure(self._ctx, ctypes.byref(c)))
        self.config = cfg

    def sizes(self) -> _abi.OitSizes:
        out = _abi.OitSizes()
        _abi.check("oit_query_sizes", self._lib.oit_query_sizes(self._ctx, ctypes.byref(out)))
        return out

    def trim(self) -> None:
        """Frees the context's frame buffers (oit_trim); configure() again before use."""
        _abi.check("oit_trim", self._lib.oit_trim(self._ctx))
        self.config = None

    def clear(self, stream=None) -> None:
        _abi.check("oit_clear", self._lib.oit_clear(self._ctx, _stream_ptr(stream)))

    def _frags(self, target, depth, rgba, count, coverage) -> _abi.OitFragments:
        if count is None:
            n = _rows(target)
            if _rows(depth) != n or _rows(rgba) != n or (n and tuple(rgba.shape[1:]) != (4,)):
                raise ValueError("fragment arrays disagree in length")
        else:
            n = int(count)
        return _abi.OitFragments(_dptr(target) if n else None, _dptr(depth) if n else None,
                                 _dptr(rgba) if n else None, n,
                                 _dptr(coverage) if (n and coverage is not None) else None)

    def draw_opaque(self, target, depth, rgba, stream=None, count=None, coverage=None) -> None:
        """The frame's opaque draw (oitRender.cpp:113-122): before any append of the frame.
        Afterwards the frame's transparent fragments are depth tested against it and the
        composite blends over its colour. Arrays as for append()."""
        f = self._frags(target, depth, rgba, count, coverage)
        _abi.check("oit_draw_opaque", self._lib.oit_draw_opaque(self._ctx, ctypes.byref(f), _stream_ptr(stream)))

    def append(self, target, depth, rgba, stream=None, count=None, coverage=None) -> None:
        """target u32 [F], depth f32 [F], rgba f32 [F, 4] premultiplied: DeviceBuffers, or raw
        device addresses with `count`; `coverage` u32 [F] sample masks with coverage shading."""
        if count is None:
            n = _rows(target)
            if _rows(depth) != n or _rows(rgba) != n or (n and tuple(rgba.shape[1:]) != (4,)):
                raise ValueError("fragment arrays disagree in length")
        else:
            n = int(count)
        f = _abi.OitFragments(_dptr(target) if n else None, _dptr(depth) if n else None,
                              _dptr(rgba) if n else None, n,
                              _dptr(coverage) if (n and coverage is not None) else None)
        _abi.check("oit_append", self._lib.oit_append(self._ctx, ctypes.byref(f), _stream_ptr(stream)))

    def resolve(self, out, stream=None) -> None:
        """out f32 [targets, 4] device tensor."""
        self._check_out(out)
        _abi.check("oit_resolve", self._lib.oit_resolve(self._ctx, _dptr(out), _stream_ptr(stream)))

    def resolve_debug(self, out, frontk_key=None, frontk_val=None, counts=None, stream=None) -> None:
        self._check_out(out)
        p = lambda t: _dptr(t) if t is not None else None  # noqa: E731
        _abi.check(
            "oit_resolve_debug",
            self._lib.oit_resolve_debug(self._ctx, _dptr(out), p(frontk_key), p(frontk_val), p(counts),
                                        _stream_ptr(stream)),
        )

    def stats(self, stream=None) -> _abi.OitStats:
        out = _abi.OitStats()
        _abi.check("oit_get_stats", self._lib.oit_get_stats(self._ctx, ctypes.byref(out), _stream_ptr(stream)))
        return out

    def selftest_codec(self, lo_bits: int, hi_bits: int) -> tuple[int, int]:
        """(mismatches, first differing bit pattern) of the device's branch-free
        sRGB8 encode against its exact table search over [lo_bits, hi_bits)."""
        bad = ctypes.c_uint64(0)
        first = ctypes.c_uint32(0)
        _abi.check("oit_selftest_codec", self._lib.oit_selftest_codec(self._ctx, lo_bits, hi_bits,
                                                                      ctypes.byref(bad), ctypes.byref(first)))
        return int(bad.value), int(first.value)

    def profile_enable(self, on: bool = True) -> None:
        _abi.check("oit_profile_enable", self._lib.oit_profile_enable(self._ctx, 1 if on else 0))

    def profile_reset(self) -> None:
        _abi.check("oit_profile_reset", self._lib.oit_profile_reset(self._ctx))

    def profile_read(self) -> dict:
        out = _abi.OitProfile()
        _abi.check("oit_profile_read", self._lib.oit_profile_read(self._ctx, ctypes.byref(out)))
        return {name: (out.ms[i], out.launches[i]) for i, name in enumerate(_abi.PROFILE_STAGE_NAMES)}

    def render(self, target, depth, rgba, out, stream=None, count=None, coverage=None) -> None:
        """One frame: clear, one append, resolve (the reference's onRender order, oitRender.cpp:43-148)."""
        self.clear(stream)
        self.append(target, depth, rgba, stream, count, coverage)
        self.resolve(out, stream)

    def finalize_srgb8(self, image, out, width, height, samples=1, supersample=1, stream=None) -> None:
        """Display image (oit_finalize_srgb8): `image` premultiplied float4
        [samples][height*supersample][width*supersample], `out` uint8 [height, width, 4]."""
        _abi.check("oit_finalize_srgb8", self._lib.oit_finalize_srgb8(self._ctx, _dptr(image), width, height, samples,
                                                                     supersample, _dptr(out), _stream_ptr(stream)))

    def _check_out(self, out) -> None:
        if self.config is None:
            raise RuntimeError("configure() first")
        if not isinstance(out, int) and tuple(out.shape) != (self.config.targets, 4):
            raise ValueError(f"out must have shape ({self.config.targets}, 4)")
