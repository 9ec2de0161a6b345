"""Runs one frame through the HIP path (C ABI via the package) and returns
numpy arrays shaped like the oracle's outputs. Used by the GPU tests only."""
from __future__ import annotations

import numpy as np


def run_gpu(algorithm, layers, tail_blend, width, height, target, depth, rgba, samples=1, ll_pool_nodes=0,
            splits=1, background=(0.5, 0.5, 0.5, 1.0), max_fragments=0, renderer=None):
    import torch

    from vk_order_independent_transparency_amd import FrameConfig, Renderer

    dev = torch.device("cuda", 0)
    T = width * height * samples
    K = layers
    own = renderer is None
    r = renderer or Renderer(0)
    try:
        cfg = FrameConfig(algorithm=algorithm, layers=layers, tail_blend=bool(tail_blend), width=width,
                          height=height, samples=samples, ll_pool_nodes=ll_pool_nodes,
                          max_fragments=max_fragments, background=background)
        r.configure(cfg)
        tg = torch.from_numpy(np.ascontiguousarray(target, dtype=np.uint32).view(np.int32)).to(dev)
        dp = torch.from_numpy(np.ascontiguousarray(depth, dtype=np.float32)).to(dev)
        cl = torch.from_numpy(np.ascontiguousarray(rgba, dtype=np.float32).reshape(-1, 4)).to(dev)
        out = torch.empty((T, 4), dtype=torch.float32, device=dev)
        fk_key = torch.empty((T, K), dtype=torch.int32, device=dev)
        fk_val = torch.empty((T, K), dtype=torch.int32, device=dev)
        counts = torch.empty((T,), dtype=torch.int32, device=dev)
        r.clear()
        F = tg.shape[0]
        bounds = np.linspace(0, F, splits + 1).astype(np.int64)
        for a, b in zip(bounds[:-1], bounds[1:]):
            r.append(tg[a:b], dp[a:b], cl[a:b])
        r.resolve_debug(out, fk_key, fk_val, counts)
        st = r.stats()
        torch.cuda.synchronize()
        return {
            "out": out.cpu().numpy(),
            "frontk_key": fk_key.cpu().numpy().view(np.uint32),
            "frontk_val": fk_val.cpu().numpy().view(np.uint32),
            "count": counts.cpu().numpy().view(np.uint32),
            "ll_overflow": int(st.ll_overflow),
            "tail_fragments": int(st.tail_fragments),
            "fragments": int(st.fragments),
        }
    finally:
        if own:
            r.close()
