"""Request loop over the OIT path (oit_server_*): continuous batching of
seeded-scene requests on one GPU with per-request A-buffer sizing from the
exact fragment count, admission against an HBM budget and resident,
reused contexts. No counterpart in the reference (an interactive frame
loop, Sample::onRender, oitRender.cpp:28-154).

The served workload: Poisson arrivals; each request a seeded scene at
1280x720, 1920x1080, 2560x1440 or 3840x2160, sphere count log-uniform in
[200, 8000], K from {8, 16, 32}, algorithm from {Linked List, Loop64,
WBOIT}, radius [0.01, 0.05] and alpha [0.1, 0.5] (the 1920x1080 workload's
ranges), tail blending on.
"""
from __future__ import annotations

import ctypes
import dataclasses
import math

import numpy as np

from . import _abi
from .scene import SceneParams, _SplitMix64

RESOLUTIONS = ((1280, 720), (1920, 1080), (2560, 1440), (3840, 2160))
ALGORITHMS = (_abi.OIT_LINKEDLIST, _abi.OIT_LOOP64, _abi.OIT_WEIGHTED)
BUDGET_96GB = 96 * 1024 ** 3


@dataclasses.dataclass(frozen=True)
class Request:
    scene: SceneParams
    algorithm: int
    layers: int
    tail_blend: bool = True

    def to_c(self, out=0, tag=0) -> _abi.OitRequest:
        r = _abi.OitRequest()
        r.scene = self.scene.to_c()
        r.algorithm = self.algorithm
        r.layers = self.layers
        r.tail_blend = 1 if self.tail_blend else 0
        r.reserved = 0
        r.out = out if isinstance(out, int) else int(out.ptr)
        r.tag = tag
        return r


def served_workload(seed: int, n: int, rate_per_s: float, resolutions=RESOLUTIONS, spheres=(200, 8000),
                    algorithms=ALGORITHMS, layers=(8, 16, 32)) -> tuple[list[Request], np.ndarray]:
    """n requests and their arrival times (microseconds from the start):
    exponential inter-arrival gaps at `rate_per_s`; every draw comes from
    `seed`, so the trace is a pure function of the arguments."""
    rng = _SplitMix64(seed ^ 0x5E4E)
    reqs, t, arrivals = [], 0.0, []
    for _ in range(n):
        w, h = rng.choice(resolutions)
        sp = rng.log_uniform_int(*spheres)
        K = rng.choice(layers)
        alg = rng.choice(algorithms)
        sseed = rng.next()
        u = rng.uniform()
        t += -math.log(1.0 - u) / rate_per_s * 1e6
        arrivals.append(t)
        reqs.append(Request(SceneParams(seed=sseed, width=w, height=h, spheres=sp, radius=(0.01, 0.05),
                                        alpha=(0.1, 0.5)), alg, K))
    return reqs, np.array(arrivals, np.float64)


def _arr(cls, items):
    a = (cls * max(len(items), 1))()
    for i, x in enumerate(items):
        a[i] = x
    return a


class Server:
    def __init__(self, device: int = 0, slots: int = 4, budget_bytes: int = BUDGET_96GB):
        self.handle = ctypes.c_void_p(0)
        cfg = _abi.OitServerConfig(slots=slots, reserved=0, budget_bytes=budget_bytes)
        _abi.check("oit_server_create", _abi.lib().oit_server_create(ctypes.byref(self.handle), device,
                                                                    ctypes.byref(cfg)))

    def submit(self, req: Request, out=0, tag: int = 0) -> None:
        r = req.to_c(out, tag)
        _abi.check("oit_server_submit", _abi.lib().oit_server_submit(self.handle, ctypes.byref(r)))

    def poll(self, max_items: int = 256) -> list[_abi.OitCompletion]:
        buf = (_abi.OitCompletion * max_items)()
        n = ctypes.c_uint32(0)
        _abi.check("oit_server_poll", _abi.lib().oit_server_poll(self.handle, buf, max_items, ctypes.byref(n)))
        return [buf[i] for i in range(n.value)]

    def drain(self, timeout_ms: int = 600000) -> None:
        _abi.check("oit_server_drain", _abi.lib().oit_server_drain(self.handle, timeout_ms))

    def replay(self, reqs: list[Request], arrivals_us, outs=None, timeout_ms: int = 600000):
        """Submits reqs[i] at arrivals_us[i] from a native driver loop and
        returns the completions in request order."""
        c_reqs = _arr(_abi.OitRequest, [r.to_c(outs[i] if outs else 0, i) for i, r in enumerate(reqs)])
        arr = np.ascontiguousarray(arrivals_us, dtype=np.float64)
        out = (_abi.OitCompletion * max(len(reqs), 1))()
        _abi.check("oit_server_replay", _abi.lib().oit_server_replay(
            self.handle, c_reqs, arr.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), len(reqs), timeout_ms, out))
        return [out[i] for i in range(len(reqs))]

    def stats(self) -> _abi.OitServerStats:
        s = _abi.OitServerStats()
        _abi.check("oit_server_get_stats", _abi.lib().oit_server_get_stats(self.handle, ctypes.byref(s)))
        return s

    def close(self) -> None:
        if self.handle:
            _abi.lib().oit_server_destroy(self.handle)
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


def latency_summary(done) -> dict:
    lat = np.array([c.t_done_us - c.t_submit_us for c in done if c.status == 0])
    if lat.size == 0:
        return {"p50_ms": None, "p99_ms": None}
    return {"p50_ms": float(np.percentile(lat, 50)) / 1e3, "p99_ms": float(np.percentile(lat, 99)) / 1e3,
            "mean_ms": float(lat.mean()) / 1e3}
