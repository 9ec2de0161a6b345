"""Synthetic sphere fragment streams (stand-in for the reference's rasterised
sphere mesh, main.cpp:334-417) and the named workloads.

The generator itself is native: csrc/oit_scene.cpp on the host and
csrc/oit_scene_gpu.hip on the device (the same stream, bit for bit), both
exported through the C ABI; this module describes workloads and moves
results into numpy or device buffers.
"""
from __future__ import annotations

import ctypes
import dataclasses

import numpy as np

from . import _abi


@dataclasses.dataclass(frozen=True)
class SceneParams:
    seed: int
    width: int
    height: int
    spheres: int
    radius: tuple[float, float]
    alpha: tuple[float, float]
    samples: int = 1
    cluster_fraction: float = 0.0
    cluster_extent: float = 0.1

    def to_c(self) -> _abi.OitSceneParams:
        return _abi.OitSceneParams(
            seed=self.seed,
            width=self.width,
            height=self.height,
            samples=self.samples,
            spheres=self.spheres,
            radius_min=self.radius[0],
            radius_max=self.radius[1],
            alpha_min=self.alpha[0],
            alpha_max=self.alpha[1],
            cluster_fraction=self.cluster_fraction,
            cluster_extent=self.cluster_extent,
        )

    @property
    def targets(self) -> int:
        return self.width * self.height * self.samples


# Committed seed list; every reported result names its seed.
SEEDS = (1, 2, 3, 4, 5, 6, 7, 8)

# 256x256, 64 spheres, r in [0.02, 0.08], alpha in [0.2, 0.6]; Linked List K=8.
SMALL_256 = SceneParams(seed=1, width=256, height=256, spheres=64, radius=(0.02, 0.08), alpha=(0.2, 0.6))
# 1920x1080, 1000 spheres, r in [0.01, 0.05], alpha in [0.1, 0.5]; K in {8, 16}.
HD_1080P = SceneParams(seed=1, width=1920, height=1080, spheres=1000, radius=(0.01, 0.05), alpha=(0.1, 0.5))
# 3840x2160, 4 samples per pixel, 4000 spheres, r in [0.005, 0.03], alpha in [0.05, 0.4].
UHD_4K_4X = SceneParams(
    seed=1, width=3840, height=2160, samples=4, spheres=4000, radius=(0.005, 0.03), alpha=(0.05, 0.4)
)


def count(p: SceneParams) -> int:
    n = ctypes.c_uint64(0)
    cp = p.to_c()
    _abi.check("oit_scene_count", _abi.lib().oit_scene_count(ctypes.byref(cp), ctypes.byref(n)))
    return int(n.value)


def generate(p: SceneParams) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """Returns (target u32[F], depth f32[F], rgba f32[F,4]) in draw order."""
    n = count(p)
    target = np.empty(max(n, 1), dtype=np.uint32)
    depth = np.empty(max(n, 1), dtype=np.float32)
    rgba = np.empty((max(n, 1), 4), dtype=np.float32)
    got = ctypes.c_uint64(0)
    cp = p.to_c()
    _abi.check(
        "oit_scene_generate_host",
        _abi.lib().oit_scene_generate_host(
            ctypes.byref(cp),
            target.ctypes.data,
            depth.ctypes.data,
            rgba.ctypes.data,
            n,
            ctypes.byref(got),
        ),
    )
    assert got.value == n
    return target[:n], depth[:n], rgba[:n]


class DeviceScene:
    """The scene's fragment stream generated on the GPU (oit_scene_gen_*).

    ``fragments`` is known after construction (the device counts once);
    ``emit`` writes the stream into device buffers on a stream and can be
    repeated. ``buffers()`` allocates a matching (target, depth, rgba) set."""

    def __init__(self, p: SceneParams, device: int = 0):
        self.params = p
        self.handle = ctypes.c_void_p(0)
        cp = p.to_c()
        _abi.check("oit_scene_gen_create", _abi.lib().oit_scene_gen_create(ctypes.byref(self.handle),
                                                                          ctypes.byref(cp), device))
        n = ctypes.c_uint64(0)
        _abi.check("oit_scene_gen_count", _abi.lib().oit_scene_gen_count(self.handle, ctypes.byref(n)))
        self.fragments = int(n.value)

    def buffers(self):
        from .device import DeviceBuffer

        n = max(self.fragments, 1)
        return DeviceBuffer((n,), np.uint32), DeviceBuffer((n,), np.float32), DeviceBuffer((n, 4), np.float32)

    def emit(self, target, depth, rgba, stream=None, capacity: int | None = None) -> None:
        from .device import stream_ptr

        cap = self.fragments if capacity is None else int(capacity)
        _abi.check("oit_scene_gen_emit", _abi.lib().oit_scene_gen_emit(
            self.handle, _ptr(target), _ptr(depth), _ptr(rgba), cap, stream_ptr(stream)))

    def close(self) -> None:
        if self.handle:
            _abi.lib().oit_scene_gen_destroy(self.handle)
            self.handle = ctypes.c_void_p(0)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _ptr(b) -> int:
    return b if isinstance(b, int) else b.ptr


class _SplitMix64:
    """The same splitmix64 the native generator uses, for workload draws."""

    def __init__(self, seed: int):
        self.s = seed & 0xFFFFFFFFFFFFFFFF

    def next(self) -> int:
        self.s = (self.s + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
        return z ^ (z >> 31)

    def uniform(self) -> float:
        return (self.next() >> 11) * 2.0 ** -53

    def choice(self, options):
        return options[self.next() % len(options)]

    def log_uniform_int(self, lo: int, hi: int) -> int:
        import math

        return int(round(math.exp(math.log(lo) + self.uniform() * (math.log(hi) - math.log(lo)))))


@dataclasses.dataclass(frozen=True)
class BatchFrame:
    scene: SceneParams
    layers: int


def batch_workload(seed: int = 1, frames: int = 64, resolutions=(256, 512, 1024, 2048)) -> list[BatchFrame]:
    """Batched workload: `frames` frames, square resolution uniform from
    `resolutions` (default {256, 512, 1024, 2048}), sphere count log-uniform
    in [100, 5000], 20 % of the spheres packed into a 0.1-wide square (local
    depth complexity up to a few hundred layers), K uniform from
    {4, 8, 16, 32}; radius [0.01, 0.05] and alpha [0.1, 0.5] as in the
    1920x1080 workload. Frame i's scene seed is drawn from `seed`, so the
    list is a pure function of (seed, frames, resolutions)."""
    rng = _SplitMix64(seed ^ 0xB47C4)
    out = []
    for _ in range(frames):
        res = rng.choice(tuple(resolutions))
        spheres = rng.log_uniform_int(100, 5000)
        K = rng.choice((4, 8, 16, 32))
        fseed = rng.next()
        out.append(BatchFrame(SceneParams(seed=fseed, width=res, height=res, spheres=spheres, radius=(0.01, 0.05),
                                          alpha=(0.1, 0.5), cluster_fraction=0.2, cluster_extent=0.1), K))
    return out
